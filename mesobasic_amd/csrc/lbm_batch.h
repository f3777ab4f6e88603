// lbm_batch.h -- device tables and launchers of the batched engine: many
// independent Shan-Chen lattices ("jobs") of different sizes and parameters
// advanced together, one launch per stencil per step.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lbm {
namespace batch {

// One job, as the kernels see it (device table entry).
struct Job {
    void* P[2];          // population buffers; the input of launch parity p is P[p ^ phase]
    float* psi;          // psi scratch (two-pass stencils)
    float* R[2];         // node shift: reference density of P[0] / P[1] (nullptr otherwise)
    const float* rho0;   // initial density for the LBfluid2D-order first step (nullptr: none)
    int nx, ny, nz;      // nz = 1 for D2Q9
    int phase;           // done & 1 when the table was uploaded (a tick boundary)
    int64_t n;           // nodes
    int64_t ps;          // population plane stride (elements)
    float tau, G, rs;    // relaxation time, coupling, storage shift
    int first;           // 1 until the first step has run (device-updated)
    int64_t remaining;   // steps still to run (device-updated, 0 = idle)
    int64_t done;        // steps run (device-updated); the current buffer is P[done & 1]
    int q;               // 9 or 19
    int blk;             // 1: D2Q9 temporal blocks (k_d2q9_block; advances by kBlkK steps per
                         // launch while more are left, then by 1: always an odd count, so
                         // P[done & 1] stays current, and the last step is a single one)
};

// Unit of work of one launch: a range of x planes (D2Q9 row kernel), of
// nodes (two-pass kernels), or a tile and a range of x planes (D3Q19 march)
// of one job.
struct Item {
    int job;
    int x0, x1;          // row kernel and D3Q19 march: planes [x0, x1)
    int64_t c0, c1;      // two-pass kernels: nodes [c0, c1)
    int y0, z0;          // D3Q19 march: tile origin
};

// D3Q19 march tile (kMarchTY rows x kMarchTZ columns of a job's y-z cross
// section); jobs with nx >= 2, ny >= kMarchTY and nz >= kMarchTZ take it.
constexpr int kMarchTY = 8, kMarchTZ = 64;

// D2Q9 temporal blocks: kBlkK steps of a kBlkT x kBlkT tile (and its halo
// of 2*kBlkK nodes) per item, in LDS.  kBlkK is odd: a launch advances a job
// by kBlkK steps while more than kBlkK are left, else by 1, and flips its
// buffer once.
constexpr int kBlkT = 32, kBlkK = 3;

// Launch parameters: job and item tables, the ticket counters of the
// persistent loop (two, alternating by launch parity) and their count.
struct Tables {
    Job* jobs;
    const Item* items;
    const int* nitems;   // device: number of valid items (tables change between graph replays)
    unsigned* tickets;   // [2]
};

// prec: LBM_FP32 / LBM_FP16S / LBM_BF16S / LBM_FP16N / LBM_BF16N (FP32
// compute in every case).  `step` is the index of the step within the tick
// (launch parity step & 1): a launch advances the jobs with remaining >
// step, whose counters (remaining, done, first) are those of the tick's
// start until advance_counters updates them after the tick's last step.
hipError_t d2q9_rows_step(int prec, const Tables& t, int step, int threads, int grid,
                          hipStream_t s);
hipError_t twopass_psi(int prec, int q, const Tables& t, int step, int grid, hipStream_t s);
hipError_t twopass_collide(int prec, int q, const Tables& t, int step, int grid, hipStream_t s);
// D3Q19 jobs in one fused pass (psi window in LDS); agrees with the two
// passes above to rounding.
hipError_t d3q19_march_step(int prec, const Tables& t, int step, int grid, hipStream_t s);
// After the `steps` single steps of a tick: remaining -= k, done += k with
// k = min(steps, remaining), and first = 0, for every job that is not on
// temporal blocks; also clears *mass_ticket, the counter of the next
// job_mass launch.  One update per tick, so the kernel classes of a tick
// depend on nothing but their own launches and run concurrently.
hipError_t advance_counters(Job* jobs, int njobs, int steps, unsigned* mass_ticket, hipStream_t s);
// D2Q9 temporal-block jobs (constant-shift forms) in tiles: each launch runs
// kBlkK steps of every block job with more than kBlkK left, 1 of the others,
// bit-identical to as many single steps of the row kernel; advance_blocks
// then updates their counters (and clears *mass_ticket too).
hipError_t d2q9_block_step(int prec, const Tables& t, int parity, int grid, hipStream_t s);
hipError_t advance_blocks(Job* jobs, int njobs, unsigned* mass_ticket, hipStream_t s);
// FP64 mass of each item's node range of its job's current populations
// (sum of the pulled f = stored/S + R over the range, as the conservation
// sums of the single-lattice forms) into part[item]; a persistent launch
// over t's items with the ticket counter *ticket (cleared by the
// advance_counters launch before it).  A non-finite value anywhere in the
// range makes the partial non-finite.
hipError_t job_mass(int prec, const Tables& t, unsigned* ticket, double* part, int grid,
                    hipStream_t s);
// Initial equilibrium of one job from rho0 (device FP32): writes the input
// buffer of the job's next launch so that its first pull yields feq.
hipError_t init_job(int prec, int q, const Job& j, int buf, hipStream_t s);
// Moments of a job's current populations into device double arrays.
hipError_t job_moments(int prec, int q, const Job& j, int buf, double* rho, double* u, double* v,
                       double* w, hipStream_t s);
hipError_t zero_u32(unsigned* p, int n, hipStream_t s);
// b[i] = a[i] for n values (FP32 to FP64).
hipError_t widen(const float* a, double* b, int64_t n, hipStream_t s);

}  // namespace batch
}  // namespace lbm
