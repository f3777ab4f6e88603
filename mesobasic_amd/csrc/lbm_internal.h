// lbm_internal.h -- types shared by the C-ABI layer and the kernel
// translation units.  Not part of the public interface.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lbm {

// Sets the calling thread's lbm_last_error() message; returns code.
int set_error(int code, const char* fmt, ...);

// Every device allocation the library makes goes through dev_malloc, so a
// test can count them and inject a failure (lbm_testing_fail_alloc).
hipError_t dev_malloc(void** p, size_t bytes);
template <typename T> inline hipError_t dev_malloc(T** p, size_t bytes)
{
    return dev_malloc(reinterpret_cast<void**>(p), bytes);
}

// Conservation sums (FP64 and low-precision forms alike): blocks of
// kConsBlock threads, each summing kConsPerThread nodes; one partial
// (mass, jx, jy, jz) per block, added on the host in block order.
constexpr int kConsBlock = 256;
constexpr int kConsPerThread = 16;
inline int64_t conservation_blocks(int64_t n)
{
    const int64_t per = static_cast<int64_t>(kConsBlock) * kConsPerThread;
    return (n + per - 1) / per;
}

// Lattice geometry on the device.  Node index c = (x*ny + y)*nz + z
// (nz = 1 in 2-D), the interior order of the reference arrays without the
// ghost shell; population n of node c lives at n*ps + c (structure of
// arrays, ps >= n rounded up and offset so that the 19 planes do not share
// HBM channel alignment).
struct Grid {
    int nx, ny, nz;
    int64_t n;
    int64_t ps;
};

// What a population buffer holds.
enum BufKind : int {
    kHoldsF = 0,   // post-streaming f (read in place)
    kHoldsPost = 1 // post-collision values (f is gathered from neighbours)
};

// Launch configuration of the fused kernel (see lbm_fused.h): tile shape
// id, x planes per workgroup, block -> tile order, non-temporal stores.
struct FusedTuning {
    int shape;
    int chunk;
    int order;
    int nts;
};

struct StepArgs {
    int q;
    int walls;
    double tau, G;
    int components = 1;   // 2: two-component Shan-Chen (mcmp2D / mcmp3D)
    double tau_b = 1.0;   // relaxation time of the second component
    // two components with walls (mcmp2DFilm / mcmp3DFilm): wall densities
    double rho_sol_a = 0.5, rho_sol_b = 0.5;
};

// FP64 path (lbm_fp64.hip).  All launches go to `s`.
hipError_t fp64_psi_from_rho(const double* rho, double* psi, int64_t n, hipStream_t s);
hipError_t fp64_psi_from_pops(const StepArgs& a, const Grid& g, const double* P, int kind,
                              double* psi, hipStream_t s);
// mode 0: moments of the gathered f; 1: rho = rho0, u = 0 (the LBfluid2D
// first step); 2: initial equilibrium (no f read; writes feq).
hipError_t fp64_collide(const StepArgs& a, const Grid& g, const double* Pin, int kind,
                        double* Pout, const double* psi, const double* rho0, int mode,
                        hipStream_t s);
// One fused pass (psi halo recomputed on chip).  Returns hipErrorNotSupported
// when the geometry does not suit it.
hipError_t fp64_fused_step(const StepArgs& a, const Grid& g, const double* Pin, int kind,
                           double* Pout, const FusedTuning& t, hipStream_t s);
bool fp64_fused_supported(const StepArgs& a, const Grid& g);
// Default fused tile shape of an FP64 lattice (-1: none fits) and whether a
// shape fits: D3Q19 tiles (lbm_fused.h) or D2Q9 row segments.
int fp64_default_shape(const Grid& g, int q);
bool fp64_shape_ok(const Grid& g, int q, int shape);
// K steps (K = fused::block_steps(shape)) of a periodic D2Q9 lattice in one
// launch from post-collision values (not the first step); hipErrorNotSupported
// where the block shape does not fit.
hipError_t fp64_block_step(const StepArgs& a, const Grid& g, const double* Pin, double* Pout, int shape,
                           hipStream_t s);
// Shan-Chen force from a psi field (FP64, reference arithmetic).
hipError_t fp64_force(const StepArgs& a, const Grid& g, const double* psi, double* fx, double* fy,
                      double* fz, hipStream_t s);
hipError_t fp64_moments(const StepArgs& a, const Grid& g, const double* P, int kind, double* rho,
                        double* u, double* v, double* w, hipStream_t s);
hipError_t fp64_get_f(const StepArgs& a, const Grid& g, const double* P, int kind, double* aos,
                      hipStream_t s);
hipError_t fp64_set_f(const StepArgs& a, const Grid& g, const double* aos, double* P,
                      hipStream_t s);
// Per-block partial sums (mass, jx, jy, jz) of the current f; returns the
// number of blocks written (4 doubles each) through nblocks.
hipError_t fp64_conservation(const StepArgs& a, const Grid& g, const double* P, int kind,
                             double* partials, int* nblocks, hipStream_t s);
int fp64_conservation_blocks(const Grid& g);

// Two components (a.components == 2, FP64): the buffer holds component A in
// planes 0..Q-1 and B in planes Q..2Q-1; psi and rho0 hold A's n values, then
// B's.  Step modes as fp64_collide.  a.walls: the film models (walls on the
// last axis, fluid-wall interaction with the wall densities).
hipError_t fp64_collide2(const StepArgs& a, const Grid& g, const double* Pin, int kind, double* Pout,
                         const double* psi, const double* rho0, int mode, hipStream_t s);
// Cross force on the component with psi_self from the other component's psi
// (other: 0 = A, 1 = B, for its wall density).
hipError_t fp64_force2(const StepArgs& a, const Grid& g, const double* psi_self, const double* psi_other,
                       int other, double* fx, double* fy, double* fz, hipStream_t s);
// Common velocity of the two components' moments, over A's u, v, w; in the
// film's wall rows the arrays keep component `own`'s velocity.
hipError_t fp64_common_velocity(const StepArgs& a, const Grid& g, const double* rA, double* uA, double* vA,
                                double* wA, const double* rB, const double* uB, const double* vB,
                                const double* wB, int own, hipStream_t s);

// Immersed-boundary capsule (capsule2D, D2Q9 FP64): device arrays of the
// IB nodes (positions, rest positions, last velocities and membrane
// forces), the fluid force arrays of the last step (n each) and the force
// the spread left on the ghost rows below and above the lattice (2*nx each:
// row y = -1, then y = ny).
struct Capsule {
    double *x, *y, *x0, *y0, *vx, *vy, *nfx, *nfy;
    double *fx, *fy, *gfx, *gfy;
    int nnodes;
    double kstiff, dt, body_fx, body_fy;
};
// Force (body + spread IB forces, stored in cap.fx/fy) and collision of the
// capsule fluid; modes as fp64_collide (mode 2: zero force, the start).
hipError_t fp64_capsule_collide(const StepArgs& a, const Grid& g, const double* Pin, int kind, double* Pout,
                                const double* rho0, int mode, const Capsule& cap, hipStream_t s);
// Ghost-row forces, membrane forces, interpolated velocities and the move of
// the IB nodes after a step that wrote P.
hipError_t fp64_capsule_finish(const StepArgs& a, const Grid& g, const double* P, int kind, const Capsule& cap,
                               hipStream_t s);

}  // namespace lbm
