/*
 * mesobasic_lbm_testing.h -- test hooks of libmesobasic_lbm.so.  These have
 * no counterpart in the reference; they expose device helpers to the parity
 * tests.
 */
#ifndef MESOBASIC_LBM_TESTING_H
#define MESOBASIC_LBM_TESTING_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* y[i] = the device exp() used by the FP64 kernels, for n host values x. */
int lbm_probe_exp(const double* x, double* y, int64_t n);

/* Host evaluation of the fused kernel's block -> tile map (block order
 * `order`, as lbm_tuning.order) for a grid of nty x ntz tiles and ncx
 * x-chunks: block b owns tile (ty[b], tz[b]) of chunk cx[b].  Arrays of
 * nty*ntz*ncx ints.  No device work. */
int lbm_probe_tile_map(int nty, int ntz, int ncx, int order, int* ty, int* tz, int* cx);

/* Streaming reference kernels over `bytes` (multiple of 16) on two device
 * buffers it allocates, run `reps` times; *ms = average time per run.
 * mode 0: copy (4 x 16 B per thread), 1: copy (8 x 16 B), 2: copy with
 * non-temporal loads/stores, 3: read only, 4: copy (1 x 16 B per thread).
 * The reference points for scaling HBM traffic counters and for the
 * achievable bandwidth. */
int lbm_probe_copy(int64_t bytes, int reps, int mode, double* ms);

/* The HIP runtime's device-to-device copy (hipMemcpyAsync) of `bytes`
 * between two buffers it allocates, `reps` times after one untimed run;
 * *ms = average time per copy.  An attainable-bandwidth figure that does not
 * depend on this library's kernels. */
int lbm_probe_memcpy(int64_t bytes, int reps, double* ms);

/* Memory skeleton of the fused D3Q19 march on an n^3 periodic lattice (n a
 * multiple of 64, at most 1024): the kernel's loads and stores (8x64 tiles,
 * its block order and x chunks, each node's 19 upstream values pulled and
 * stored unchanged) without its arithmetic, LDS window, ring or barriers,
 * on two buffers it allocates, `reps` times after two untimed runs;
 * *ms = average time per run.  elem_bytes 2, 4 or 8; layout 0: one plane per
 * population (the lattice handles' layout), 1: the 19 populations of each
 * 64-node z segment stored together.  The attainable rate of that access
 * pattern, against which the step kernels are measured. */
int lbm_probe_march(int n, int elem_bytes, int layout, int reps, double* ms);

/* Device allocations the library has made so far: every hipMalloc of the
 * lattice handles and the batch engine goes through one wrapper that counts
 * it (the probes above allocate on their own and are not counted). */
int64_t lbm_testing_alloc_count(void);

/* Injected allocation failure: the next `after` device allocations succeed
 * and the one after them fails as hipErrorOutOfMemory would (once, then the
 * hook disarms); a negative value disarms it.  Lets a test force the
 * allocation-failure paths of lbm_create and lbm_batch_submit. */
void lbm_testing_fail_alloc(int64_t after);

#ifdef __cplusplus
}
#endif
#endif
