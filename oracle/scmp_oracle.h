/*
 * scmp_oracle.h -- CPU restatement of mesoBasic's Shan-Chen single-component
 * multiphase (SCMP) lattice-Boltzmann step, used ONLY as a test checker and
 * as the timed CPU baseline.  Nothing in the product path links this file.
 *
 * It follows the reference arithmetic statement by statement, in the
 * reference's own array-of-structures layout with a one-node ghost shell:
 *   src/lattice_boltzmann/LBUtils/Stencil.cpp:5-120      (tables, order)
 *   src/lattice_boltzmann/LBUtils/LBfluid3D.cpp:180-347  (feq, macros,
 *                                                          collide, stream,
 *                                                          z bounce-back)
 *   src/lattice_boltzmann/LBUtils/LBfluid3D.cpp:430-543  (ghost shells)
 *   src/lattice_boltzmann/LBUtils/LBfluid2D.cpp:201-388  (D2Q9 engine)
 *   src/lattice_boltzmann/LBUtils/LBfluid2D.cpp:495-580  (D2Q9 ghosts)
 *   src/lattice_boltzmann/LBTypes/scmp3D.cpp:54-190      (model, psi, force)
 *   src/lattice_boltzmann/LBTypes/mcmp2D.cpp:97-201      (2-D step order)
 * Indices are 64-bit (the reference uses int, which overflows at 512^3 on
 * one rank); the x-slab split over threads replaces the MPI slab split of
 * src/lattice_boltzmann/LBApp.cpp:46-60 and leaves results unchanged because
 * every per-node operation reads only the previous phase's arrays.
 *
 * Parity status: pinned bit-for-bit against the compiled reference sources
 * (oracle/_ref harness, fixtures in tests/golden) and by closed-form checks
 * (stencil moments, uniform fixed point, mass/momentum conservation,
 * periodic-shift invariance, D3Q19 NZ=1 vs D2Q9).  psi() calls the host
 * libm exp(), exactly as scmp3D.cpp:189 does.
 */
#ifndef SCMP_ORACLE_H
#define SCMP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct scmp_oracle scmp_oracle;

/* q = 9 (D2Q9, nz must be 1) or 19 (D3Q19).
 * order = 3: step = macros -> ghost rho -> force -> collide -> stream
 *            (scmp3D::updateLatticeBoltzmann, scmp3D.cpp:90-119)
 * order = 2: step = force -> collide -> stream -> walls -> macros -> ghost rho
 *            (mcmp2D::updateLatticeBoltzmann + LBfluid2D::updateFluid_SC,
 *             mcmp2D.cpp:97-123, LBfluid2D.cpp:227-312)
 * walls = 1: halfway bounce-back walls on the last axis (z in 3-D,
 *            LBfluid3D.cpp:316-347; y in 2-D, LBfluid2D.cpp:363-388).
 * nthreads: x-slab worker threads (1 = serial, as one MPI rank). */
scmp_oracle* scmp_oracle_create(int q, int nx, int ny, int nz, double tau,
                                double G, int order, int walls, int nthreads);
void scmp_oracle_destroy(scmp_oracle* h);

/* Initial density, interior nodes, reference order (x slowest, last axis
 * fastest).  Velocities are set to zero, as scmp3D.cpp:67-75 does. */
void scmp_oracle_set_rho(scmp_oracle* h, const double* rho0);

/* scmp3D::initLatticeBoltzmann minus the rand() fill (scmp3D.cpp:76-81):
 * ghost rho, Shan-Chen force, f = feq(rho, u + tau*F/rho). */
void scmp_oracle_init(scmp_oracle* h);

void scmp_oracle_step(scmp_oracle* h, int nsteps);

/* Copies the reference arrays, interior only, reference order.  rho/u/v/w
 * are the r,u,v,w member arrays (their time level depends on the step
 * order); f is [node][q].  Any pointer may be NULL.  w is ignored in 2-D. */
void scmp_oracle_get(const scmp_oracle* h, double* rho, double* u, double* v,
                     double* w, double* f);
/* Shan-Chen force arrays fx, fy, fz (interior, reference order). */
void scmp_oracle_get_force(const scmp_oracle* h, double* fx, double* fy,
                           double* fz);
/* Moments of the current f (rho = sum f, u = sum f e / rho), computed with
 * the reference's summation order. */
void scmp_oracle_moments(const scmp_oracle* h, double* rho, double* u,
                         double* v, double* w);

/* Overwrites f (interior, [node][q]) and sets the r,u,v,w arrays to its
 * moments; used to start from a snapshot. */
void scmp_oracle_set_f(scmp_oracle* h, const double* f);

/* Replaces the exp() used inside psi(); NULL restores libm exp.  Only for
 * sensitivity experiments (how far a 1-ulp change in psi travels). */
void scmp_oracle_set_exp(scmp_oracle* h, double (*fn)(double));

int64_t scmp_oracle_nodes(const scmp_oracle* h);

/* Seeded inputs.  noise: rho = mean + amp*(x/4294967295.0 - 0.5) with x the
 * successive outputs of std::mt19937(seed) in reference loop order.
 * droplet: tanh profile of radius R and width W between rl and rv, centred
 * at ((N-1)/2) per axis.  mt19937: raw outputs (known-answer tests). */
void scmp_oracle_noise_rho(uint32_t seed, int64_t n, double mean, double amp,
                           double* out);
void scmp_oracle_mt19937(uint32_t seed, int64_t n, uint32_t* out);
void scmp_oracle_droplet_rho(int nx, int ny, int nz, double R, double W,
                             double rl, double rv, double* out);

/* ---- two components (mcmp2D / mcmp3D) -------------------------------------
 * Components A and B with their own tau and one cross coupling G (the
 * reference's "LBApp/G"); order 2 is mcmp2D's step (forces and common
 * velocity, then LBfluid2D::updateFluid_SC of A and of B), order 3 mcmp3D's
 * (macros of both, forces and common velocity, collide-stream of both).
 * comp: 0 = A, 1 = B.  The arrays have the reference's meanings: after a
 * step of order 3 the u arrays hold the common velocity, in order 2 each
 * component's own velocity. */
typedef struct mcmp_oracle mcmp_oracle;
mcmp_oracle* mcmp_oracle_create(int q, int nx, int ny, int nz, double tauA, double tauB, double G,
                                int order, int nthreads);
void mcmp_oracle_destroy(mcmp_oracle* m);
/* The film models mcmp2DFilm / mcmp3DFilm: walls on the last axis (rows 1
 * and n), fluid-wall interaction through psi(rhoA_sol / rhoB_sol), bounce-back
 * of both components.  Call before init. */
void mcmp_oracle_set_film(mcmp_oracle* m, double rhoA_sol, double rhoB_sol);
void mcmp_oracle_set_rho(mcmp_oracle* m, const double* rhoA, const double* rhoB);
void mcmp_oracle_init(mcmp_oracle* m);
void mcmp_oracle_step(mcmp_oracle* m, int nsteps);
void mcmp_oracle_get(const mcmp_oracle* m, int comp, double* rho, double* u, double* v, double* w,
                     double* f);
void mcmp_oracle_get_force(const mcmp_oracle* m, int comp, double* fx, double* fy, double* fz);
void mcmp_oracle_moments(const mcmp_oracle* m, int comp, double* rho, double* u, double* v,
                         double* w);

/* ---- immersed-boundary capsule (capsule2D) ---------------------------------
 * One D2Q9 fluid of relaxation time tau in the LBfluid2D step order with
 * bounce-back walls in y and no Shan-Chen force; nnodes IB nodes tied to
 * their rest positions by springs of stiffness kstiff, spread onto the fluid
 * and moved with the interpolated velocity (capsule2D.cpp:84-150,
 * IBcapsules2D.cpp:130-484).  body_fx, body_fy: the body force on the rows
 * between the walls (capsule2D.cpp:128-136 uses 0.000008, 0). */
typedef struct capsule_oracle capsule_oracle;
capsule_oracle* capsule_oracle_create(int nx, int ny, double tau, int nnodes, double kstiff,
                                      double dt, double body_fx, double body_fy);
void capsule_oracle_destroy(capsule_oracle* c);
void capsule_oracle_set_rho(capsule_oracle* c, const double* rho0);
/* Node positions, which are also the rest positions. */
void capsule_oracle_set_nodes(capsule_oracle* c, const double* x, const double* y);
/* Node positions only (rest positions kept): a displaced start. */
void capsule_oracle_move_nodes(capsule_oracle* c, const double* x, const double* y);
/* IBcapsules2D::initCircularCapsule(xo, yo, rad). */
void capsule_oracle_circle(capsule_oracle* c, double xo, double yo, double rad);
void capsule_oracle_init(capsule_oracle* c);
void capsule_oracle_step(capsule_oracle* c, int nsteps);
/* The fluid's r,u,v arrays, f ([node][9]) and force arrays (interior). */
void capsule_oracle_get(const capsule_oracle* c, double* rho, double* u, double* v, double* f,
                        double* fx, double* fy);
/* Node positions, velocities (last interpolation) and membrane forces. */
void capsule_oracle_get_nodes(const capsule_oracle* c, double* x, double* y, double* vx, double* vy,
                              double* fx, double* fy);
/* Force accumulated on the ghost rows below and above the lattice (2*nx
 * values each: row j = 0, then row j = ny+1). */
void capsule_oracle_get_ghost_force(const capsule_oracle* c, double* gfx, double* gfy);

#ifdef __cplusplus
}
#endif
#endif
