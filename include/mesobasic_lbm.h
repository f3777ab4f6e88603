/*
 * mesobasic_lbm.h -- C ABI of the MI355X Shan-Chen lattice-Boltzmann step.
 *
 * One handle = one lattice resident in HBM (structure-of-arrays populations,
 * A/B buffers).  Plain C types only; host arrays are interior nodes in the
 * reference's order (x slowest, last axis fastest: index (x*NY + y)*NZ + z,
 * or x*NY + y in 2-D), caller-owned.  Every call returns 0 on success or a
 * negative LBM_E* code; lbm_last_error() gives the message.  A handle is not
 * thread-safe; distinct handles are independent.
 *
 * Each entry point replaces a piece of the reference's LBfluid3D/LBfluid2D
 * and scmp3D interface (flowzario/mesoBasic, src/lattice_boltzmann); the
 * reference's C++ classes are what a maintainer binds these to (the drop-in
 * classes in mesobasic_amd/host do exactly that).
 */
#ifndef MESOBASIC_LBM_H
#define MESOBASIC_LBM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBM_ABI_VERSION 2

enum lbm_status {
    LBM_OK = 0,
    LBM_EINVAL = -1,   /* bad argument or description */
    LBM_ENOMEM = -2,   /* device allocation failed or over the HBM budget */
    LBM_EHIP = -3,     /* HIP runtime error */
    LBM_ESTATE = -4,   /* call out of order (e.g. step before init) */
    LBM_ENODEV = -5,   /* no usable gfx950 device */
    LBM_EJOB = -6      /* a batch job failed (its mass total went non-finite) */
};

enum lbm_precision {
    LBM_FP64 = 0,      /* FP64 compute and storage: the drop-in parity path */
    LBM_FP32 = 1,      /* FP32 compute and storage */
    LBM_FP16S = 2,     /* FP32 compute, FP16 storage of f_i - w_i*rho_shift */
    LBM_BF16S = 3,     /* FP32 compute, BF16 storage of f_i - w_i*rho_shift */
    /* Node-shifted 16-bit storage: f_i - w_i*R(x), where R(x) is the density
     * that the node x reading the value had after the previous step (kept
     * as an FP32 field, 4 B per node, double-buffered).  Resolves densities
     * far from any single shift (droplets, walls); rho_shift is unused. */
    LBM_FP16N = 4,     /* FP32 compute, FP16 node-shifted storage */
    LBM_BF16N = 5      /* FP32 compute, BF16 node-shifted storage */
};

/* AUTO: the fused kernel where it measured faster (D3Q19: at least 2^21
 * nodes, and in the FP32/16-bit forms a tile that divides the lattice other
 * than 8x32; FP64 also with partial tiles where no tile divides it;
 * D2Q9: at least 2^18 nodes, or ny = 64, 128 or 256), else the two passes.
 * FUSED asks for it wherever a shape fits (D3Q19 partial tiles: ny >= 4,
 * nz >= 32; FP64 results unchanged to the bit). */
enum lbm_kernel {
    LBM_KERNEL_AUTO = 0,
    LBM_KERNEL_TWO_PASS = 1, /* density pass, then force+collide+stream pass */
    LBM_KERNEL_FUSED = 2     /* one pass, psi halo recomputed in LDS */
};

/* Lattice description.  Mirrors the keys the reference reads:
 *   q          "LBApp/stencil" D2Q9 -> 9, D3Q19 -> 19 (scmp3D.cpp:30-32)
 *   nx,ny,nz   "Domain/nx,ny,nz" (LBApp.cpp:18-20); nz = 1 for q = 9
 *   tau        "LBApp/tauA" (scmp3D.cpp:24-25)
 *   G          "LBApp/G" (scmp3D.cpp:21)
 *   order      3: scmp3D step order (scmp3D.cpp:90-119)
 *              2: LBfluid2D step order (LBfluid2D.cpp:227-312)
 *   walls      1: halfway bounce-back walls on the last axis
 *              (LBfluid3D.cpp:316-347 in 3-D, LBfluid2D.cpp:363-388 in 2-D)
 *   precision  enum lbm_precision
 *   rho_shift  density whose equilibrium is subtracted before 16-bit storage
 *   kernel     enum lbm_kernel
 *   components 0 or 1: one component (scmp3D, and its D2Q9 form);
 *              2: two-component Shan-Chen, "LBApp/type" mcmp2D (q = 9,
 *              order 2, mcmp2D.cpp:97-123) or mcmp3D (q = 19, order 3,
 *              mcmp3D.cpp:97-123): G couples the components ("LBApp/G",
 *              mcmp2D.cpp:20), tau is component A's ("LBApp/tauA"),
 *              tau_b component B's ("LBApp/tauB", mcmp2D.cpp:26).  FP64,
 *              two-pass kernels.  With walls = 1: the film models
 *              mcmp2DFilm / mcmp3DFilm (walls on the last axis; the wall
 *              rows carry no force and no common velocity, and fluid nodes
 *              next to them interact with psi(rho_sol_a / rho_sol_b),
 *              mcmp2DFilm.cpp:149-219, mcmp3DFilm.cpp:167-253).
 *   tau_b      relaxation time of component B (components = 2)
 *   rho_sol_a, rho_sol_b   wall densities of A and B ("LBApp/rhoA_sol",
 *              "rhoB_sol", mcmp2DFilm.cpp:21-22; components = 2, walls = 1)
 */
typedef struct lbm_desc {
    int q;
    int nx, ny, nz;
    double tau;
    double G;
    int order;
    int walls;
    int precision;
    double rho_shift;
    int kernel;
    int components;
    double tau_b;
    double rho_sol_a, rho_sol_b;
} lbm_desc;

typedef struct lbm_handle lbm_handle;

/* Allocates the lattice on the current device.  Replaces the LBfluid3D /
 * LBfluid2D constructors and allocateFs (LBfluid3D.cpp:26-95,
 * LBfluid2D.cpp:26-89). */
int lbm_create(const lbm_desc* desc, lbm_handle** out);
void lbm_destroy(lbm_handle* h);

/* Initial density (host, nx*ny*nz doubles); velocities zero.  Replaces the
 * setRho/setU/setV/setW fill of scmp3D::initLatticeBoltzmann
 * (scmp3D.cpp:63-76). */
int lbm_set_rho(lbm_handle* h, const double* rho0);

/* ghost rho -> Shan-Chen force -> f = feq(rho, u + tau*F/rho).  Replaces
 * scmp3D.cpp:78-80 (ghostNodesRho, calculateShanChenForces, setFtoFeq). */
int lbm_init_equilibrium(lbm_handle* h);

/* Advances nsteps time steps on the handle's stream (asynchronous).
 * Replaces scmp3D::updateLatticeBoltzmann (scmp3D.cpp:90-119) or, in order
 * 2, force + LBfluid2D::updateFluid_SC (mcmp2D.cpp:111-121), including the
 * bounce-back walls when enabled. */
int lbm_step(lbm_handle* h, int nsteps);

/* As lbm_step, bracketed by HIP events on the handle's stream; waits and
 * returns the elapsed milliseconds. */
int lbm_step_timed(lbm_handle* h, int nsteps, double* ms);

enum lbm_fields_kind {
    LBM_FIELDS_REFERENCE = 0, /* the reference's r,u,v,w member arrays:
                                 order 3 -> moments of f(N-1),
                                 order 2 -> moments of f(N); rho0, 0 at N=0 */
    LBM_FIELDS_CURRENT = 1    /* moments of the current f(N) */
};

/* Density and bare velocity (LBfluid3D::macros, LBfluid3D.cpp:208-244),
 * host arrays of nx*ny*nz doubles; any pointer may be NULL; w ignored in
 * 2-D.  Synchronous. */
int lbm_get_fields(lbm_handle* h, int kind, double* rho, double* u, double* v,
                   double* w);

/* Shan-Chen force held in the reference's fx,fy,fz arrays after the last
 * completed step: F(psi(rho)) of the density the step started from
 * (scmp3D.cpp:141-179; rho0 before the first step).  Host arrays of
 * nx*ny*nz doubles; fz ignored in 2-D.  Synchronous.  Used for the
 * Shan-Chen physical velocity u* = u + F/(2 rho) (LBfluid2D.cpp:164-172). */
int lbm_get_force(lbm_handle* h, double* fx, double* fy, double* fz);

/* Post-streaming populations f (after walls), host [node][q] doubles as the
 * reference's f array (LBfluid3D.hpp:29).  Synchronous. */
int lbm_get_f(lbm_handle* h, double* f);

/* Replaces the populations (host [node][q]); the next step starts from them
 * with the reference r,u arrays taken as their moments. */
int lbm_set_f(lbm_handle* h, const double* f);

/* FP64 totals over the lattice of the current f: out[0] = mass,
 * out[1..3] = momentum (x, y, z; z = 0 in 2-D).  Synchronous. */
int lbm_conservation(lbm_handle* h, double out[4]);

/* ---- components -------------------------------------------------------------
 * Two-component handles (lbm_desc.components = 2) hold components A (comp 0)
 * and B (comp 1); the calls above act on A, these on either.  Replacing the
 * reference's fA / fB members of mcmp2D / mcmp3D (mcmp2D.hpp:14-16). */

/* Components of the handle (1 or 2). */
int lbm_components(const lbm_handle* h);

/* Initial density of one component (fA.setRho / fB.setRho with zero
 * velocity, mcmp2D.cpp:67-81); lbm_set_rho sets component A. */
int lbm_set_rho_component(lbm_handle* h, int comp, const double* rho0);

/* The reference's r,u,v,w member arrays of one component
 * (LBM_FIELDS_REFERENCE), or the moments of its current f
 * (LBM_FIELDS_CURRENT).  Two components: in order 3 (mcmp3D) the u arrays
 * after a step hold the common velocity of both components
 * (mcmp3D.cpp:201-212), and r the density of f(N-1); in order 2 (mcmp2D)
 * each component's own moments of f(N).  Synchronous. */
int lbm_get_fields_component(lbm_handle* h, int comp, int kind, double* rho, double* u, double* v,
                             double* w);

/* The fx,fy,fz arrays of one component: with two components the cross force
 * ((-G) psi_c(x)) sum (w e) psi_other(x+e) (mcmp2D.cpp:179-182). */
int lbm_get_force_component(lbm_handle* h, int comp, double* fx, double* fy, double* fz);

/* Populations of one component, as lbm_get_f / lbm_set_f (set both
 * components before the next step). */
int lbm_get_f_component(lbm_handle* h, int comp, double* f);
int lbm_set_f_component(lbm_handle* h, int comp, const double* f);

/* FP64 totals of one component's current f, as lbm_conservation. */
int lbm_conservation_component(lbm_handle* h, int comp, double out[4]);

/* ---- immersed-boundary capsule ----------------------------------------------
 * The reference's capsule2D model (capsule2D.cpp:84-150 with IBcapsules2D,
 * IBcapsules2D.cpp:130-484): a D2Q9 fluid in the LBfluid2D step order (a
 * handle with q = 9, order 2, one component, FP64; walls = 1 gives the
 * model's bounce-back walls in y) whose force arrays hold, in place of the
 * Shan-Chen force (G is not used), a body force on the rows between the
 * walls plus the membrane forces of nnodes IB nodes, F = -kstiff (x - x0),
 * spread with the bilinear kernel; after each fluid step the nodes move with
 * the interpolated velocity u* = u + F/(2 rho), x += u* dt.
 * lbm_get_force returns the force arrays of the last step (zero at the
 * start).  Nodes with a non-finite x, or x outside [-1, nx), are skipped as
 * the reference skips them; y must stay in [-1, ny) (beyond it the
 * reference indexes outside its ghost shell; such nodes are skipped here). */
typedef struct lbm_capsule_desc {
    int nnodes;        /* IB nodes ("LBApp/nnodes", IBcapsules2D.cpp:46) */
    double kstiff;     /* membrane stiffness ("LBApp/kstiff", :48) */
    double dt;         /* time step of the node move ("Time/dt", :25) */
    double body_fx;    /* body force (capsule2D.cpp:132-133: 0.000008, 0) */
    double body_fy;
} lbm_capsule_desc;

/* Attaches a capsule to a handle before lbm_init_equilibrium; the start is
 * then f = feq(rho0, 0) with zero forces (capsule2D.cpp:92-118). */
int lbm_capsule_attach(lbm_handle* h, const lbm_capsule_desc* d);
/* Node positions (host, nnodes each), which are also the rest positions x0
 * (IBcapsules2D::initCircularCapsule sets both, IBcapsules2D.cpp:130-141).
 * Required before the first step. */
int lbm_capsule_set_nodes(lbm_handle* h, const double* x, const double* y);
/* Node positions, the velocities of the last interpolation and the membrane
 * forces of the last step (host arrays of nnodes doubles; any may be NULL). */
int lbm_capsule_get_nodes(lbm_handle* h, double* x, double* y, double* vx, double* vy, double* fx, double* fy);
/* IB nodes of the attached capsule (0: none). */
int lbm_capsule_nodes(const lbm_handle* h);

/* ---- asynchronous snapshot -----------------------------------------------------
 * The output step of the reference models (scmp3D::outputLatticeBoltzmann,
 * scmp3D.cpp:185-245; mcmp2D.cpp:255-300) reads r, u and fx of the lattice
 * and writes a VTK file while the run waits.  lbm_snapshot_begin enqueues the
 * same fields on the handle's stream (as lbm_get_fields_component with kind,
 * then lbm_get_force_component, of component comp) into a device buffer and
 * their copy to pinned host memory on a second stream of the handle, and
 * returns at once: the following lbm_step calls run while the copy goes on.
 * lbm_snapshot_wait waits for that copy and hands out the seven arrays of
 * nx*ny*nz doubles (any pointer may be NULL) and the step count they were
 * taken at.  One snapshot is held at a time: a second begin first waits for
 * the previous copy (whose data it then overwrites). */
int lbm_snapshot_begin(lbm_handle* h, int comp, int kind);
int lbm_snapshot_wait(lbm_handle* h, int64_t* step, double* rho, double* u, double* v, double* w,
                      double* fx, double* fy, double* fz);

/* Number of completed steps since lbm_init_equilibrium / lbm_set_f. */
int64_t lbm_steps_done(const lbm_handle* h);

/* Stream the handle launches on (a hipStream_t); NULL restores the handle's
 * own stream.  The caller keeps ownership of a stream it passes. */
int lbm_set_stream(lbm_handle* h, void* stream);
void* lbm_get_stream(const lbm_handle* h);
int lbm_synchronize(lbm_handle* h);

/* Launch configuration of the fused kernel.  Results do not depend on it
 * (the FP64 path stays bit-identical to the reference); only speed does.
 * No reference counterpart: the reference has no launch configuration.
 *   shape     tile of the y-z cross-section per workgroup: 0: 1x512, 1: 2x256,
 *             2: 1x256 (whole z rows, nz must equal 512 / 256 / 256),
 *             3: 8x64, 4: 4x64, 5: 8x32 (also partial: any ny >= TY,
 *             nz >= TZ); two nodes per thread (16-bit
 *             storage at tau = 1 without walls, nz a multiple of 128):
 *             6: 8x128, 7: 4x128; D2Q9 FP64 (any ny): a segment of 8: 256,
 *             9: 128, 10: 64 nodes of a y row; D2Q9 FP64 temporal
 *             blocks (periodic, one component; K steps per launch on T x T
 *             tiles, T dividing nx and ny, the last step of each lbm_step
 *             call a single step): 11: 16x16 x 2 steps, 12: 16x16 x 4,
 *             13: 8x8 x 4; -1 = the defaults (a pair tile where one
 *             applies; for D2Q9 the narrowest segment that covers a row of
 *             <= 128 nodes, else 256, and a temporal block for periodic
 *             lattices of up to 2^20 nodes that one fits).  Every fused
 *             shape needs nx >= 2.
 *   chunk     x planes (D2Q9: x rows) marched per workgroup
 *   order     block -> tile order: 0 plain, 1 contiguous runs per XCD,
 *             2 y-tiles of one column per XCD, 3 a y-band x all z-tiles
 *             per XCD (D2Q9 segments: one fixed order, the value is kept)
 *   nt_store  1 = stores carry the non-temporal hint */
typedef struct lbm_tuning {
    int shape;
    int chunk;
    int order;
    int nt_store;
} lbm_tuning;
int lbm_set_tuning(lbm_handle* h, const lbm_tuning* t);
int lbm_get_tuning(const lbm_handle* h, lbm_tuning* t);

/* Device bytes held by the handle, and the kernel it resolved to. */
int64_t lbm_device_bytes(const lbm_handle* h);
int lbm_kernel_in_use(const lbm_handle* h);

/* Message for the last failing call on this thread. */
const char* lbm_last_error(void);

/* Seeded inputs (host only, no device work).  noise: rho = mean +
 * amp*(x/4294967295.0 - 0.5) for successive std::mt19937(seed) outputs x in
 * reference loop order -- the scmp3D.cpp:67-69 fill with a reproducible
 * generator.  droplet: (rl+rv)/2 - (rl-rv)/2*tanh(2(r-R)/W), r from the
 * geometric centre ((N-1)/2 per axis).  nz = 1 for 2-D. */
int lbm_fill_noise(uint32_t seed, int64_t n, double mean, double amp, double* out);
int lbm_fill_droplet(int nx, int ny, int nz, double R, double W, double rl,
                     double rv, double* out);

int lbm_abi_version(void);

/* ---- batched and serving form -------------------------------------------
 * Many independent Shan-Chen lattices ("jobs") of different sizes,
 * stencils and parameters advanced together: one persistent launch per
 * stencil class per step over device-resident job tables, a fixed number of
 * steps per tick captured in a hipGraph, jobs admitted against an HBM
 * budget and retired as they finish (continuous batching).  FP32 arithmetic
 * with FP32, FP16 or BF16 storage of f - w*rho_shift.  A job's arithmetic
 * does not depend on the other jobs in the batch.  No counterpart in the
 * reference, whose driver runs one lattice per process (MesoExecute.cpp:
 * 107-152); each job follows the reference step order of its stencil
 * (D2Q9: LBfluid2D.cpp:227-312, D3Q19: scmp3D.cpp:90-119). */

typedef struct lbm_batch lbm_batch;

typedef struct lbm_batch_desc {
    int precision;          /* LBM_FP32, LBM_FP16S, LBM_BF16S, LBM_FP16N or LBM_BF16N */
    double rho_shift;       /* storage shift (0 is used for LBM_FP32) */
    int64_t hbm_budget;     /* bytes the jobs may hold; 0 = 90% of free HBM */
    int steps_per_tick;     /* steps per tick (rounded up to even) */
    int use_graph;          /* 1 = replay each tick as a captured hipGraph */
    int autotune;           /* 1 = time launch configurations on the first ticks */
} lbm_batch_desc;

typedef struct lbm_job_desc {
    int q;                  /* 9 or 19 */
    int nx, ny, nz;         /* nz = 1 for q = 9 */
    double tau, G;
    int64_t steps;          /* steps to run; the job finishes after them */
} lbm_job_desc;

int lbm_batch_create(const lbm_batch_desc* desc, lbm_batch** out);
void lbm_batch_destroy(lbm_batch* b);

/* Device bytes a job would hold. */
int64_t lbm_batch_job_bytes(const lbm_batch* b, const lbm_job_desc* job);

/* Admits a job with initial density rho0 (host, nx*ny*nz doubles, zero
 * velocity).  Returns LBM_ENOMEM, with nothing allocated, if the job does
 * not fit the remaining budget: the caller keeps it queued and retries
 * after jobs finish.  The job starts with the next tick.  A job holds fewer
 * than 2^30 nodes (LBM_EINVAL otherwise; larger lattices belong on a
 * single-lattice handle). */
int lbm_batch_submit(lbm_batch* b, const lbm_job_desc* job, const double* rho0, int64_t* job_id);

/* Runs one tick (steps_per_tick steps of every unfinished job; a job stops
 * exactly at its step count), waits for it, and reports how many jobs are
 * finished and not yet collected.  The tick ends with an FP64 mass total of
 * every job it advanced (the last launch of the tick graph): a job whose
 * total is not finite -- a NaN or Inf anywhere in its populations, from a
 * poisoned initial density or a diverging run -- is retired there as
 * failed, takes no further steps and no further share of the launches, and
 * leaves the other jobs' results unchanged. */
int lbm_batch_tick(lbm_batch* b, int* n_finished);

/* lbm_batch_tick in two halves, so that the ticks of several batches run on
 * the device together: begin uploads the tables and launches the tick (with
 * the read-back of its counters and mass totals) on the batch's stream and
 * returns; end waits for it and does the rest of lbm_batch_tick (retiring
 * failed jobs, reporting finished ones).  Between the two, calls that change
 * the batch (submit, collect, read, extend, cancel, the set_* calls and
 * another begin) return LBM_ESTATE.  No counterpart in the reference (the
 * batched form is this library's own). */
int lbm_batch_tick_begin(lbm_batch* b);
int lbm_batch_tick_end(lbm_batch* b, int* n_finished);

/* Ids of finished, uncollected jobs (up to max), failed ones included. */
int lbm_batch_finished(lbm_batch* b, int64_t* ids, int max, int* n);

/* Density and velocity (moments of the final populations, FP64) of a
 * finished job into host arrays (any may be NULL); frees the job.  For a
 * failed job the arrays are filled the same way (with its non-finite
 * values), the job is freed, and the call returns LBM_EJOB with the step of
 * the failure in lbm_last_error(). */
int lbm_batch_collect(lbm_batch* b, int64_t job_id, double* rho, double* u, double* v, double* w);

/* The reference's r,u,v,w and fx,fy,fz arrays of an admitted job after the
 * steps it has run (status.steps_done), as lbm_get_fields
 * (LBM_FIELDS_REFERENCE) and lbm_get_force return them for a single lattice
 * in the job's step order (D3Q19: scmp3D, D2Q9: LBfluid2D): the moments,
 * in FP64, of the job's stored populations, and the FP64 force of the density
 * its last step started from.  Host arrays of nx*ny*nz doubles, any may be
 * NULL; fz is 0 for D2Q9.  The job stays admitted.  Call between ticks.  The
 * output step of the reference's models (scmp3D.cpp:127-133,
 * mcmp2D.cpp:131-137) for a job of the batch. */
int lbm_batch_read(lbm_batch* b, int64_t job_id, double* rho, double* u, double* v, double* w, double* fx,
                   double* fy, double* fz);

/* Gives a finished (not failed) job nsteps more steps: it runs again from
 * the next tick, from the populations it finished with, exactly as if it had
 * been submitted with the larger step count.  With lbm_batch_read this runs a
 * job to an output cadence (MesoExecute.cpp:107-152: output every
 * nstep/numOutputs steps).  LBM_ESTATE for a running job, LBM_EJOB for a
 * failed one. */
int lbm_batch_extend(lbm_batch* b, int64_t job_id, int64_t nsteps);

enum lbm_job_state {
    LBM_JOB_RUNNING = 0,   /* admitted, steps left */
    LBM_JOB_DONE = 1,      /* ran all its steps; waiting to be collected */
    LBM_JOB_FAILED = 2     /* retired: its mass total went non-finite */
};

typedef struct lbm_job_status {
    int state;              /* enum lbm_job_state */
    int64_t steps_done;     /* steps run (for a failed job: at the tick that saw the failure) */
    int64_t steps_finite;   /* steps after which the last mass check was finite */
    double mass;            /* FP64 total mass at the last check (NaN before the first tick) */
} lbm_job_status;

/* State of an admitted, uncollected job. */
int lbm_batch_job_status(const lbm_batch* b, int64_t job_id, lbm_job_status* st);

/* Cancels an admitted job, running or finished: frees its device memory at
 * once (ticks are synchronous, so none of its work is in flight) and
 * removes it from the work tables before the next tick.  The other jobs'
 * results do not change. */
int lbm_batch_cancel(lbm_batch* b, int64_t job_id);

/* Jobs admitted and not collected; device bytes they hold; steps run. */
int lbm_batch_active(const lbm_batch* b);
int64_t lbm_batch_bytes_in_use(const lbm_batch* b);
int64_t lbm_batch_steps_run(const lbm_batch* b);

/* Scheduling: with max_nodes > 0, a tick advances only the unfinished jobs
 * with the least remaining work (nodes x steps left), taken in that order
 * while their node count stays within max_nodes (always at least one job);
 * the others stay resident and wait.  Shortest-remaining-work-first lowers
 * the latency of small requests when more work is admitted than the GPU
 * needs to stay busy.  0 (default): every unfinished job, every tick.
 * Results do not depend on the schedule. */
int lbm_batch_set_work_limit(lbm_batch* b, int64_t max_nodes);

/* Bound on waiting under the work limit: a resident, unfinished job that the
 * last `ticks` ticks did not advance goes ahead of the shorter ones (longest
 * wait first) until it is advanced, so a large job is not starved by a steady
 * stream of small ones.  0 (default): no bound.  Results do not depend on
 * it. */
int lbm_batch_set_max_wait(lbm_batch* b, int ticks);

/* Concurrent kernel classes (1, default): within a tick, the steps of each
 * kernel class (a D2Q9 row width, a two-pass stencil, the D3Q19 march, the
 * temporal blocks) run on a stream of their own, forked from the batch's
 * stream and joined before the tick's mass check, so a class of a few small
 * jobs overlaps the large ones.  0: every launch on the batch's stream, one
 * class after another.  Results do not depend on it (no reference
 * counterpart: the reference runs one lattice per process). */
int lbm_batch_set_concurrency(lbm_batch* b, int on);

/* Launch configuration the batch uses (after autotuning): D2Q9 rows per
 * work item (at most; a row-width class whose items would not fill two
 * rounds of resident workgroups gets shorter items, down to 2 rows) and
 * whether D2Q9 jobs use the row-marching kernel (1), the two-pass kernels
 * (0) or temporal blocks (2; see lbm_batch_set_launch).  The tuner times
 * each on one tick. */
int lbm_batch_tuning(const lbm_batch* b, int* d2q9_chunk, int* d2q9_rows_kernel);

/* Sets the launch configuration (and stops the tuner): D2Q9 rows per work
 * item (>= 1; at most, as above), D2Q9 jobs on the row-marching kernel (1), the two-pass
 * kernels (0) or temporal blocks (2: 3 steps per launch on 32x32 tiles held in
 * LDS, single steps for the last few; the constant-shift forms -- the
 * node-shift forms keep the row kernel), and the kernel of D3Q19 jobs: 0 the two-pass kernels, 2 the
 * fused march kernel for every job with nx >= 2, ny >= 8 and nz >= 64, 1
 * (default) the march for such jobs in the node-shift forms when their y-z
 * cross-section fills at least 90% of the 8 x 64 march tiles it needs, and
 * the two passes for the others.  A job's D3Q19 kernel depends only on its shape and
 * this mode.  D2Q9 results
 * do not depend on the configuration; D3Q19 results of the two kernels
 * agree to FP32 rounding, not bit for bit, and the tuner never changes the
 * D3Q19 mode.
 * lbm_batch_get_launch reports the current configuration.  No reference
 * counterpart (the reference runs one lattice per process). */
int lbm_batch_set_launch(lbm_batch* b, int d2q9_chunk, int d2q9_rows_kernel, int d3q19_march);
int lbm_batch_get_launch(const lbm_batch* b, int* d2q9_chunk, int* d2q9_rows_kernel,
                         int* d3q19_march);

/* ---- allocation-free entry point ------------------------------------------
 * lbm_conservation_noalloc: the totals of lbm_conservation, equal to them bit
 * for bit, computed without any device allocation.  No counterpart in the
 * reference; it is for callers that cannot allocate device memory (code
 * running between HIP graph captures, tight serving loops, harnesses that
 * forbid hipMalloc).  Its block partial sums live in a fixed device array of
 * 2 MiB + 32 B per GPU, held until the process ends, so it takes lattices of
 * at most 2^28 nodes (LBM_EINVAL beyond).  The device current at the call
 * must be the handle's (LBM_ESTATE otherwise).  Every call on a GPU uses that
 * GPU's one array: calls that overlapped would race on it, so the library
 * serialises the calls of a process with a lock held until the totals are on
 * the host.  Synchronous. */
int lbm_conservation_noalloc(lbm_handle* h, double out[4]);

#ifdef __cplusplus
}
#endif
#endif
