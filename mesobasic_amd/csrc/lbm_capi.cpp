// lbm_capi.cpp -- the C ABI declared in include/mesobasic_lbm.h: handle
// lifetime, device memory, step sequencing and host<->device copies.  All
// arithmetic runs in the HIP kernels; there is no host fallback.
#include "../../include/mesobasic_lbm.h"

#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "lbm_fused.h"
#include "lbm_handle.h"
#include "lbm_internal.h"
#include "lbm_lowp.h"

namespace {
thread_local std::string g_err;
std::atomic<int64_t> g_alloc_count{0};   // device allocations made (test hook)
std::atomic<int64_t> g_fail_after{-1};   // allocations let through before one fails; -1: none
}  // namespace

namespace lbm {
// Records the message returned by lbm_last_error() on this thread.
int set_error(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

hipError_t dev_malloc(void** p, size_t bytes)
{
    *p = nullptr;
    // injected failure (lbm_testing_fail_alloc): counts down, fails once at 0
    int64_t k = g_fail_after.load();
    while (k >= 0 && !g_fail_after.compare_exchange_weak(k, k - 1)) {
    }
    if (k == 0) return hipErrorOutOfMemory;
    g_alloc_count.fetch_add(1);
    return hipMalloc(p, bytes);
}
}  // namespace lbm

namespace {

#define fail lbm::set_error

#define HIP_OK(expr)                                                                  \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess)                                                         \
            return fail(LBM_EHIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_),      \
                        __FILE__, __LINE__);                                          \
    } while (0)

// Plane stride: rounded up to 256 B and skewed by 512 B per population so
// that the Q planes of a large power-of-two lattice do not start on the same
// HBM channel.
int64_t plane_stride(int64_t n, int elem_bytes)
{
    const int64_t per256 = 256 / elem_bytes;
    int64_t ps = (n + per256 - 1) / per256 * per256;
    if ((ps * elem_bytes) % 4096 == 0) ps += 512 / elem_bytes;
    return ps;
}

}  // namespace

namespace {

int ensure_tmp(lbm_handle* h, int64_t bytes)
{
    if (h->tmp_bytes >= bytes) return LBM_OK;
    if (h->tmp) hipFree(h->tmp);
    h->tmp = nullptr;
    h->tmp_bytes = 0;
    HIP_OK(lbm::dev_malloc(&h->tmp, bytes));
    h->tmp_bytes = bytes;
    return LBM_OK;
}

// Components of a handle, and the populations of component comp in buffer b
// (component c holds planes c*q .. c*q+q-1).
int comps(const lbm_handle* h) { return h->a.components; }
double* pops64(lbm_handle* h, int b, int comp)
{
    return static_cast<double*>(h->buf[b]) + static_cast<int64_t>(comp) * h->d.q * h->g.ps;
}

int check_comp(const lbm_handle* h, int comp)
{
    if (comp < 0 || comp >= comps(h))
        return fail(LBM_EINVAL, "component %d out of range (handle has %d)", comp, comps(h));
    return LBM_OK;
}

// Two-component step (mcmp2D / mcmp3D): psi of both components, then one
// pass that forms the cross forces and common velocity and collides both.
int run_step2(lbm_handle* h)
{
    const int src = h->cur, dst = 1 - h->cur;
    const int mode = h->first_step_stored ? 1 : 0;
    const int64_t n = h->g.n;
    for (int c = 0; c < 2; c++) {
        if (mode == 1) HIP_OK(lbm::fp64_psi_from_rho(h->rho0 + c * n, h->psi + c * n, n, h->stream));
        else HIP_OK(lbm::fp64_psi_from_pops(h->a, h->g, pops64(h, src, c), h->kind[src], h->psi + c * n,
                                            h->stream));
    }
    HIP_OK(lbm::fp64_collide2(h->a, h->g, static_cast<const double*>(h->buf[src]), h->kind[src],
                              static_cast<double*>(h->buf[dst]), h->psi, h->rho0, mode, h->stream));
    h->kind[dst] = lbm::kHoldsPost;
    h->cur = dst;
    h->first_step_stored = false;
    h->prev_valid = true;
    h->steps++;
    return LBM_OK;
}

// One capsule2D step: force (body + spread IB forces) and collision, then
// the IB nodes' velocities and move (capsule2D.cpp:123-150).
int run_step_capsule(lbm_handle* h)
{
    if (!h->capsule_nodes_set) return fail(LBM_ESTATE, "capsule nodes not set (lbm_capsule_set_nodes)");
    const int src = h->cur, dst = 1 - h->cur;
    auto* Pin = static_cast<const double*>(h->buf[src]);
    auto* Pout = static_cast<double*>(h->buf[dst]);
    const int mode = h->first_step_stored ? 1 : 0;
    HIP_OK(lbm::fp64_capsule_collide(h->a, h->g, Pin, h->kind[src], Pout, h->rho0, mode, h->cap, h->stream));
    HIP_OK(lbm::fp64_capsule_finish(h->a, h->g, Pout, lbm::kHoldsPost, h->cap, h->stream));
    h->kind[dst] = lbm::kHoldsPost;
    h->cur = dst;
    h->first_step_stored = false;
    h->prev_valid = true;
    h->steps++;
    return LBM_OK;
}

int run_step(lbm_handle* h)
{
    if (comps(h) == 2) return run_step2(h);
    if (h->has_capsule) return run_step_capsule(h);
    const int src = h->cur, dst = 1 - h->cur;
    if (!is_lowp(h)) {
        auto* Pin = static_cast<const double*>(h->buf[src]);
        auto* Pout = static_cast<double*>(h->buf[dst]);
        const int mode = h->first_step_stored ? 1 : 0;
        if (h->kernel == LBM_KERNEL_FUSED && mode == 0) {
            HIP_OK(lbm::fp64_fused_step(h->a, h->g, Pin, h->kind[src], Pout, h->tune, h->stream));
        } else {
            if (!h->psi) return fail(LBM_ESTATE, "psi field not allocated");
            if (mode == 1) HIP_OK(lbm::fp64_psi_from_rho(h->rho0, h->psi, h->g.n, h->stream));
            else HIP_OK(lbm::fp64_psi_from_pops(h->a, h->g, Pin, h->kind[src], h->psi, h->stream));
            HIP_OK(lbm::fp64_collide(h->a, h->g, Pin, h->kind[src], Pout, h->psi, h->rho0, mode,
                                     h->stream));
        }
    } else {
        if (!h->psi) return fail(LBM_ESTATE, "psi field not allocated");
        HIP_OK(lbm::lowp_step(h->d.precision, h->a, h->g, h->d.rho_shift, h->buf[src],
                              h->kind[src], h->buf[dst], h->first_step_stored ? h->rho0 : nullptr,
                              reinterpret_cast<float*>(h->psi), h->rref[src], h->rref[dst],
                              h->kernel, h->tune, h->stream));
    }
    h->kind[dst] = lbm::kHoldsPost;
    h->cur = dst;
    h->first_step_stored = false;
    h->prev_valid = true;
    h->steps++;
    return LBM_OK;
}

// Enqueues the moments (rho, u, v, w) of component comp of buffer b into the
// device arrays d, d+n, d+2n, d+3n.
int moments_to_device(lbm_handle* h, int b, int comp, double* d)
{
    const int64_t n = h->g.n;
    if (!is_lowp(h))
        HIP_OK(lbm::fp64_moments(h->a, h->g, pops64(h, b, comp), h->kind[b], d, d + n, d + 2 * n,
                                 d + 3 * n, h->stream));
    else
        HIP_OK(lbm::lowp_moments(h->d.precision, h->a, h->g, h->d.rho_shift, h->buf[b],
                                 h->kind[b], h->rref[b], d, d + n, d + 2 * n, d + 3 * n,
                                 h->stream));
    return LBM_OK;
}

// Copies k consecutive device arrays of n doubles to the non-NULL host
// arrays and waits.
int arrays_to_host(lbm_handle* h, const double* d, double* const* outs, int k)
{
    const int64_t n = h->g.n;
    for (int i = 0; i < k; i++)
        if (outs[i])
            HIP_OK(hipMemcpyAsync(outs[i], d + i * n, n * 8, hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    return LBM_OK;
}

// Moments of component comp of buffer b into the four host arrays.
int moments_to_host(lbm_handle* h, int b, double* rho, double* u, double* v, double* w, int comp = 0)
{
    int rc = ensure_tmp(h, 4 * h->g.n * 8);
    if (rc) return rc;
    rc = moments_to_device(h, b, comp, h->tmp);
    if (rc) return rc;
    double* outs[4] = {rho, u, v, w};
    return arrays_to_host(h, h->tmp, outs, 4);
}

int check_desc(const lbm_desc* d)
{
    if (!d) return fail(LBM_EINVAL, "lbm_create: desc is NULL");
    if (d->q != 9 && d->q != 19) return fail(LBM_EINVAL, "q must be 9 or 19 (got %d)", d->q);
    if (d->nx < 1 || d->ny < 1 || d->nz < 1)
        return fail(LBM_EINVAL, "lattice sizes must be >= 1 (got %d x %d x %d)", d->nx, d->ny, d->nz);
    if (d->q == 9 && d->nz != 1) return fail(LBM_EINVAL, "D2Q9 needs nz == 1 (got %d)", d->nz);
    if (d->order != 2 && d->order != 3)
        return fail(LBM_EINVAL, "order must be 2 (LBfluid2D) or 3 (scmp3D) (got %d)", d->order);
    if (!(d->tau > 0.0) || !std::isfinite(d->tau)) return fail(LBM_EINVAL, "tau must be > 0");
    if (!std::isfinite(d->G)) return fail(LBM_EINVAL, "G must be finite");
    if (d->walls) {
        const int last = d->q == 19 ? d->nz : d->ny;
        if (last < 3) return fail(LBM_EINVAL, "walls need >= 3 nodes on the wall axis (got %d)", last);
    }
    if (d->precision < LBM_FP64 || d->precision > LBM_BF16N)
        return fail(LBM_EINVAL, "unknown precision %d", d->precision);
    if (d->kernel < LBM_KERNEL_AUTO || d->kernel > LBM_KERNEL_FUSED)
        return fail(LBM_EINVAL, "unknown kernel %d", d->kernel);
    const int64_t n = static_cast<int64_t>(d->nx) * d->ny * d->nz;
    if (n >= (int64_t(1) << 32)) return fail(LBM_EINVAL, "lattice too large (%lld nodes)", (long long)n);
    if (d->components < 0 || d->components > 2)
        return fail(LBM_EINVAL, "components must be 0, 1 or 2 (got %d)", d->components);
    if (d->components == 2) {
        if (d->precision != LBM_FP64)
            return fail(LBM_EINVAL, "two components run in FP64 only");
        if (d->walls && !(std::isfinite(d->rho_sol_a) && std::isfinite(d->rho_sol_b)))
            return fail(LBM_EINVAL, "two components with walls: rho_sol_a, rho_sol_b must be finite");
        if (d->kernel == LBM_KERNEL_FUSED)
            return fail(LBM_EINVAL, "two components run on the two-pass kernels");
        if (!(d->tau_b > 0.0) || !std::isfinite(d->tau_b)) return fail(LBM_EINVAL, "tau_b must be > 0");
    }
    return LBM_OK;
}

}  // namespace

extern "C" {

int lbm_abi_version(void) { return LBM_ABI_VERSION; }

const char* lbm_last_error(void) { return g_err.c_str(); }

int lbm_create(const lbm_desc* desc, lbm_handle** out)
{
    if (!out) return fail(LBM_EINVAL, "lbm_create: out is NULL");
    *out = nullptr;
    int rc = check_desc(desc);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(LBM_ENODEV, "lbm_create: no HIP device");
    auto* h = new lbm_handle();
    h->d = *desc;
    if (hipGetDevice(&h->device) != hipSuccess) {
        delete h;
        return fail(LBM_ENODEV, "lbm_create: no current HIP device");
    }
    h->elem = (desc->precision == LBM_FP64) ? 8 : (desc->precision == LBM_FP32 ? 4 : 2);
    h->g.nx = desc->nx; h->g.ny = desc->ny; h->g.nz = desc->nz;
    h->g.n = static_cast<int64_t>(desc->nx) * desc->ny * desc->nz;
    h->g.ps = plane_stride(h->g.n, h->elem);
    h->a.q = desc->q; h->a.walls = desc->walls ? 1 : 0;
    h->a.tau = desc->tau; h->a.G = desc->G;
    h->a.components = desc->components == 2 ? 2 : 1;
    h->a.tau_b = desc->components == 2 ? desc->tau_b : desc->tau;
    h->a.rho_sol_a = desc->rho_sol_a;
    h->a.rho_sol_b = desc->rho_sol_b;
    const int nc = h->a.components;
    if (desc->precision == LBM_FP64) {
        const bool fused_ok = lbm::fp64_fused_supported(h->a, h->g);
        if (desc->kernel == LBM_KERNEL_FUSED && !fused_ok) {
            delete h;
            return fail(LBM_EINVAL, "fused FP64 kernel does not support this lattice");
        }
        // automatic choice: D3Q19 where fused::auto_fused19 holds, partial
        // tiles included (FP64, GLUPS, fused vs two-pass: 200^3 13.2 vs 11.1,
        // 250^3 13.7 vs 10.9, 300^3 12.8 vs 11.0; 100^3 10.3 vs 11.8 is
        // below the size threshold); D2Q9 small lattices stay on the two
        // passes unless their rows are full segments
        const bool fused_auto =
            fused_ok && (desc->q == 19 ? lbm::fused::auto_fused19(h->g, lbm::fp64_default_shape(h->g, desc->q), true)
                                       : (h->g.n >= lbm::fused::kRowAutoMinNodes ||
                                          lbm::fused::row_full(h->g, lbm::fp64_default_shape(h->g, desc->q))));
        h->kernel = desc->kernel == LBM_KERNEL_AUTO
                        ? (fused_auto && nc == 1 ? LBM_KERNEL_FUSED : LBM_KERNEL_TWO_PASS)
                        : desc->kernel;
        // temporal blocks (K steps per launch) for the small periodic D2Q9
        // lattices, which are latency-bound one step at a time
        const int blk = nc == 1 ? lbm::fused::default_block(h->g, desc->q, desc->walls != 0) : -1;
        if (blk >= 0 && fused_ok && desc->kernel != LBM_KERNEL_TWO_PASS) {
            h->kernel = LBM_KERNEL_FUSED;
            h->block = blk;
        }
    } else {
        h->kernel = lbm::lowp_resolve_kernel(desc->precision, h->a, h->g, desc->kernel);
        if (h->kernel < 0) {
            delete h;
            return fail(LBM_EINVAL, "requested kernel does not support this lattice/precision");
        }
    }
    h->tune.shape = desc->precision == LBM_FP64
                        ? lbm::fp64_default_shape(h->g, desc->q)
                        : lbm::lowp_default_shape(desc->precision, h->a, h->g);
    // x planes per workgroup: at most 256 for the FP64 and the pair tiles
    // (fewer prologue planes; 128 measured 1.5-2.3% faster than 64 at 256^3
    // and 512^3 on MI355X, and 256 another 0.3-0.9% at 512^3), 64 for the
    // others (flat between 32 and 256), fewer where the lattice would not
    // fill the CUs (fused::default_chunk)
    const bool wide = desc->precision == LBM_FP64 || lbm::fused::is_pair_shape(h->tune.shape);
    const int max_chunk = wide ? 256 : 64;
    // resident waves per CU: 2 per SIMD for FP64, the pair tiles and the
    // one-node 16-bit tiles with walls, 4 for the other FP32-compute tiles
    const int waves_per_cu = (wide || desc->walls) ? 8 : 16;
    h->tune.chunk = lbm::fused::default_chunk(h->g, h->tune.shape, max_chunk, waves_per_cu);
    if (lbm::fused::is_row_shape(h->tune.shape))
        h->tune.chunk = lbm::fused::default_row_chunk(h->g, h->tune.shape);
    const int64_t pop_bytes = h->g.ps * desc->q * h->elem * nc;
    auto cleanup = [&](int code) {
        lbm_destroy(h);
        return code;
    };
    for (int b = 0; b < 2; b++) {
        hipError_t e = lbm::dev_malloc(&h->buf[b], pop_bytes);
        if (e != hipSuccess)
            return cleanup(fail(LBM_ENOMEM, "hipMalloc(%lld B) for populations: %s",
                                (long long)pop_bytes, hipGetErrorString(e)));
    }
    h->bytes = 2 * pop_bytes;
    {
        // psi field: the two-pass kernels use it every step, and every
        // handle runs them for initialisation and for the LBfluid2D-order
        // first step, so it is always allocated.
        if (lbm::dev_malloc(&h->psi, nc * h->g.n * 8) != hipSuccess)
            return cleanup(fail(LBM_ENOMEM, "hipMalloc for psi failed"));
        h->bytes += nc * h->g.n * 8;
    }
    if (lbm::dev_malloc(&h->rho0, nc * h->g.n * 8) != hipSuccess)
        return cleanup(fail(LBM_ENOMEM, "hipMalloc for rho0 failed"));
    h->bytes += nc * h->g.n * 8;
    if (lbm::lowp_node_shift(desc->precision)) {
        for (int b = 0; b < 2; b++)
            if (lbm::dev_malloc(&h->rref[b], h->g.n * 4) != hipSuccess)
                return cleanup(fail(LBM_ENOMEM, "hipMalloc for the node-shift density failed"));
        h->bytes += 2 * h->g.n * 4;
    }
    if (hipStreamCreateWithFlags(&h->own, hipStreamNonBlocking) != hipSuccess)
        return cleanup(fail(LBM_EHIP, "hipStreamCreate failed"));
    h->stream = h->own;
    if (hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess)
        return cleanup(fail(LBM_EHIP, "hipEventCreate failed"));
    *out = h;
    return LBM_OK;
}

void lbm_destroy(lbm_handle* h)
{
    if (!h) return;
    if (h->stream) hipStreamSynchronize(h->stream);
    for (void* b : h->buf)
        if (b) hipFree(b);
    if (h->psi) hipFree(h->psi);
    for (float* r : h->rref)
        if (r) hipFree(r);
    if (h->rho0) hipFree(h->rho0);
    if (h->tmp) hipFree(h->tmp);
    if (h->ev0) hipEventDestroy(h->ev0);
    if (h->ev1) hipEventDestroy(h->ev1);
    if (h->copy_stream) {
        hipStreamSynchronize(h->copy_stream);
        hipStreamDestroy(h->copy_stream);
    }
    if (h->ev_snap) hipEventDestroy(h->ev_snap);
    if (h->ev_copied) hipEventDestroy(h->ev_copied);
    if (h->snap_dev) hipFree(h->snap_dev);
    if (h->snap_host) hipHostFree(h->snap_host);
    if (h->capsule_mem) hipFree(h->capsule_mem);
    if (h->own) hipStreamDestroy(h->own);
    delete h;
}

int lbm_set_rho(lbm_handle* h, const double* rho0) { return lbm_set_rho_component(h, 0, rho0); }

int lbm_set_rho_component(lbm_handle* h, int comp, const double* rho0)
{
    if (!h || !rho0) return fail(LBM_EINVAL, "lbm_set_rho: NULL argument");
    int rc = check_comp(h, comp);
    if (rc) return rc;
    HIP_OK(hipMemcpyAsync(h->rho0 + comp * h->g.n, rho0, h->g.n * 8, hipMemcpyHostToDevice, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    h->initialised = false;
    return LBM_OK;
}

int lbm_components(const lbm_handle* h) { return h ? h->a.components : 0; }

int lbm_init_equilibrium(lbm_handle* h)
{
    if (!h) return fail(LBM_EINVAL, "lbm_init_equilibrium: NULL handle");
    if (!h->psi) return fail(LBM_ESTATE, "psi field not allocated");
    if (comps(h) == 2) {
        for (int c = 0; c < 2; c++)
            HIP_OK(lbm::fp64_psi_from_rho(h->rho0 + c * h->g.n, h->psi + c * h->g.n, h->g.n, h->stream));
        HIP_OK(lbm::fp64_collide2(h->a, h->g, nullptr, lbm::kHoldsF, static_cast<double*>(h->buf[0]),
                                  h->psi, h->rho0, 2, h->stream));
    } else if (!is_lowp(h)) {
        double* psi = h->psi;
        HIP_OK(lbm::fp64_psi_from_rho(h->rho0, psi, h->g.n, h->stream));
        HIP_OK(lbm::fp64_collide(h->a, h->g, nullptr, lbm::kHoldsF, static_cast<double*>(h->buf[0]),
                                 psi, h->rho0, 2, h->stream));
    } else {
        HIP_OK(lbm::lowp_init(h->d.precision, h->a, h->g, h->d.rho_shift, h->rho0, h->buf[0],
                              reinterpret_cast<float*>(h->psi), h->rref[0], h->stream));
    }
    HIP_OK(hipStreamSynchronize(h->stream));
    if (h->has_capsule) {
        // the capsule model's start: zero force arrays (capsule2D.cpp:104-105)
        HIP_OK(hipMemsetAsync(h->cap.fx, 0, 2 * h->g.n * 8, h->stream));
        HIP_OK(lbm::fp64_capsule_collide(h->a, h->g, nullptr, lbm::kHoldsF, static_cast<double*>(h->buf[0]),
                                         h->rho0, 2, h->cap, h->stream));
        HIP_OK(hipStreamSynchronize(h->stream));
    }
    h->cur = 0;
    h->kind[0] = lbm::kHoldsF;
    h->steps = 0;
    h->initialised = true;
    h->first_step_stored = (h->d.order == 2);
    h->prev_valid = false;
    h->ref_is_rho0 = true;
    return LBM_OK;
}

int lbm_step(lbm_handle* h, int nsteps)
{
    if (!h) return fail(LBM_EINVAL, "lbm_step: NULL handle");
    if (!h->initialised) return fail(LBM_ESTATE, "lbm_step before lbm_init_equilibrium / lbm_set_f");
    if (nsteps < 0) return fail(LBM_EINVAL, "lbm_step: nsteps < 0");
    const int K = lbm::fused::block_steps(h->block);
    for (int s = 0; s < nsteps;) {
        // K steps per launch where a temporal block is set, after the first
        // step (post-collision values in); the last step of the call is a
        // single one, so the other buffer holds f(N-1) as the reference
        // arrays and the force getters expect
        if (h->block >= 0 && !h->first_step_stored && h->kind[h->cur] == lbm::kHoldsPost &&
            nsteps - s > K) {
            const int src = h->cur, dst = 1 - h->cur;
            HIP_OK(lbm::fp64_block_step(h->a, h->g, static_cast<const double*>(h->buf[src]),
                                        static_cast<double*>(h->buf[dst]), h->block, h->stream));
            h->kind[dst] = lbm::kHoldsPost;
            h->cur = dst;
            h->prev_valid = false;
            h->steps += K;
            s += K;
            continue;
        }
        int rc = run_step(h);
        if (rc) return rc;
        s++;
    }
    return LBM_OK;
}

int lbm_step_timed(lbm_handle* h, int nsteps, double* ms)
{
    if (!h || !ms) return fail(LBM_EINVAL, "lbm_step_timed: NULL argument");
    HIP_OK(hipEventRecord(h->ev0, h->stream));
    int rc = lbm_step(h, nsteps);
    if (rc) return rc;
    HIP_OK(hipEventRecord(h->ev1, h->stream));
    HIP_OK(hipEventSynchronize(h->ev1));
    float t = 0.f;
    HIP_OK(hipEventElapsedTime(&t, h->ev0, h->ev1));
    *ms = t;
    return LBM_OK;
}

int lbm_get_fields(lbm_handle* h, int kind, double* rho, double* u, double* v, double* w)
{
    return lbm_get_fields_component(h, 0, kind, rho, u, v, w);
}

}  // extern "C"

namespace {

// Enqueues component comp's r,u,v,w arrays (kind as lbm_get_fields) into the
// device arrays out, out+n, out+2n, out+3n.  Uses the staging buffer (8n
// doubles at most), which must not overlap out.
int fields_to_device(lbm_handle* h, int comp, int kind, double* out)
{
    const int64_t n = h->g.n;
    if (kind == LBM_FIELDS_CURRENT) return moments_to_device(h, h->cur, comp, out);
    if (kind != LBM_FIELDS_REFERENCE) return fail(LBM_EINVAL, "unknown fields kind %d", kind);
    if (h->steps == 0 && h->ref_is_rho0) {
        // r = rho0, u = v = w = 0 (scmp3D.cpp:63-76 before any step; with two
        // components the common velocity initLatticeBoltzmann leaves, which is
        // 0: mcmp2D.cpp:85, mcmp3D.cpp:89)
        HIP_OK(hipMemcpyAsync(out, h->rho0 + comp * n, n * 8, hipMemcpyDeviceToDevice, h->stream));
        HIP_OK(hipMemsetAsync(out + n, 0, 3 * n * 8, h->stream));
        return LBM_OK;
    }
    if (h->d.order == 2 || h->steps == 0) return moments_to_device(h, h->cur, comp, out);
    if (h->steps == 1 && h->d.order == 3 && !h->prev_valid)
        return fail(LBM_ESTATE, "previous populations unavailable");
    const int prev = 1 - h->cur;
    if (comps(h) == 1)
        // order 3: the arrays hold the moments computed at the start of step N
        return moments_to_device(h, prev, 0, out);
    // mcmp3D: r = this component's density of f(N-1), u = the common
    // velocity formed from both components' moments of f(N-1)
    int rc = ensure_tmp(h, 8 * n * 8);
    if (rc) return rc;
    double* A = h->tmp;
    double* B = h->tmp + 4 * n;
    HIP_OK(lbm::fp64_moments(h->a, h->g, pops64(h, prev, 0), h->kind[prev], A, A + n, A + 2 * n, A + 3 * n,
                             h->stream));
    HIP_OK(lbm::fp64_moments(h->a, h->g, pops64(h, prev, 1), h->kind[prev], B, B + n, B + 2 * n, B + 3 * n,
                             h->stream));
    HIP_OK(hipMemcpyAsync(out, comp == 0 ? A : B, n * 8, hipMemcpyDeviceToDevice, h->stream));
    HIP_OK(lbm::fp64_common_velocity(h->a, h->g, A, A + n, A + 2 * n, A + 3 * n, B, B + n, B + 2 * n,
                                     B + 3 * n, comp, h->stream));
    HIP_OK(hipMemcpyAsync(out + n, A + n, 3 * n * 8, hipMemcpyDeviceToDevice, h->stream));
    return LBM_OK;
}

// Enqueues component comp's fx, fy, fz arrays into out, out+n, out+2n
// (staging buffer: (5*components + 3) n doubles, not overlapping out).
int force_to_device(lbm_handle* h, int comp, double* out)
{
    const int64_t n = h->g.n;
    const int nc = comps(h);
    if (h->has_capsule) {
        // the capsule's force arrays: body force plus the spread IB forces of
        // the last step (zero at the start)
        HIP_OK(hipMemcpyAsync(out, h->cap.fx, n * 8, hipMemcpyDeviceToDevice, h->stream));
        HIP_OK(hipMemcpyAsync(out + n, h->cap.fy, n * 8, hipMemcpyDeviceToDevice, h->stream));
        HIP_OK(hipMemsetAsync(out + 2 * n, 0, n * 8, h->stream));
        return LBM_OK;
    }
    // per component: density the last step started from, then psi; then F
    int rc = ensure_tmp(h, (5 * nc + 3) * n * 8);
    if (rc) return rc;
    double* F = h->tmp + 5 * nc * n;
    const bool from_rho0 = h->ref_is_rho0 && (h->steps == 0 || (h->steps == 1 && h->d.order == 2));
    for (int c = 0; c < nc; c++) {
        double* rho = h->tmp + 5 * c * n;
        double* psi = rho + 4 * n;
        if (from_rho0) {
            HIP_OK(hipMemcpyAsync(rho, h->rho0 + c * n, n * 8, hipMemcpyDeviceToDevice, h->stream));
        } else {
            const int b = (h->steps == 0) ? h->cur : 1 - h->cur;
            rc = moments_to_device(h, b, c, rho);
            if (rc) return rc;
        }
        HIP_OK(lbm::fp64_psi_from_rho(rho, psi, n, h->stream));
    }
    if (nc == 1) {
        HIP_OK(lbm::fp64_force(h->a, h->g, h->tmp + 4 * n, F, F + n, F + 2 * n, h->stream));
    } else {
        const double* ps = h->tmp + (5 * comp + 4) * n;
        const double* po = h->tmp + (5 * (1 - comp) + 4) * n;
        HIP_OK(lbm::fp64_force2(h->a, h->g, ps, po, 1 - comp, F, F + n, F + 2 * n, h->stream));
    }
    HIP_OK(hipMemcpyAsync(out, F, 3 * n * 8, hipMemcpyDeviceToDevice, h->stream));
    return LBM_OK;
}

// Device buffer of the snapshot (7n doubles) and its pinned host copy.
int ensure_snapshot(lbm_handle* h)
{
    const int64_t bytes = 7 * h->g.n * 8;
    if (h->snap_dev) return LBM_OK;
    HIP_OK(lbm::dev_malloc(&h->snap_dev, bytes));
    HIP_OK(hipHostMalloc(&h->snap_host, bytes, hipHostMallocDefault));
    HIP_OK(hipStreamCreateWithFlags(&h->copy_stream, hipStreamNonBlocking));
    HIP_OK(hipEventCreateWithFlags(&h->ev_snap, hipEventDisableTiming));
    HIP_OK(hipEventCreateWithFlags(&h->ev_copied, hipEventDisableTiming));
    return LBM_OK;
}

}  // namespace

extern "C" {

int lbm_get_fields_component(lbm_handle* h, int comp, int kind, double* rho, double* u, double* v,
                             double* w)
{
    if (!h) return fail(LBM_EINVAL, "lbm_get_fields: NULL handle");
    int rc = check_comp(h, comp);
    if (rc) return rc;
    if (!h->initialised) return fail(LBM_ESTATE, "lbm_get_fields before initialisation");
    // the fields are formed behind the staging area the two-component
    // reference arrays use (8n doubles), then copied out
    const int64_t n = h->g.n;
    rc = ensure_tmp(h, 12 * n * 8);
    if (rc) return rc;
    double* out = h->tmp + 8 * n;
    rc = fields_to_device(h, comp, kind, out);
    if (rc) return rc;
    double* outs[4] = {rho, u, v, w};
    return arrays_to_host(h, out, outs, 4);
}

int lbm_snapshot_begin(lbm_handle* h, int comp, int kind)
{
    if (!h) return fail(LBM_EINVAL, "lbm_snapshot_begin: NULL handle");
    int rc = check_comp(h, comp);
    if (rc) return rc;
    if (!h->initialised) return fail(LBM_ESTATE, "lbm_snapshot_begin before initialisation");
    rc = ensure_snapshot(h);
    if (rc) return rc;
    // the previous snapshot's copy still reads snap_dev and fills snap_host
    if (h->snap_pending) HIP_OK(hipEventSynchronize(h->ev_copied));
    h->snap_pending = false;
    rc = fields_to_device(h, comp, kind, h->snap_dev);
    if (rc) return rc;
    rc = force_to_device(h, comp, h->snap_dev + 4 * h->g.n);
    if (rc) return rc;
    HIP_OK(hipEventRecord(h->ev_snap, h->stream));
    HIP_OK(hipStreamWaitEvent(h->copy_stream, h->ev_snap, 0));
    HIP_OK(hipMemcpyAsync(h->snap_host, h->snap_dev, 7 * h->g.n * 8, hipMemcpyDeviceToHost, h->copy_stream));
    HIP_OK(hipEventRecord(h->ev_copied, h->copy_stream));
    h->snap_pending = true;
    h->snap_step = h->steps;
    return LBM_OK;
}

int lbm_snapshot_wait(lbm_handle* h, int64_t* step, double* rho, double* u, double* v, double* w,
                      double* fx, double* fy, double* fz)
{
    if (!h) return fail(LBM_EINVAL, "lbm_snapshot_wait: NULL handle");
    if (!h->snap_pending) return fail(LBM_ESTATE, "lbm_snapshot_wait without lbm_snapshot_begin");
    HIP_OK(hipEventSynchronize(h->ev_copied));
    h->snap_pending = false;
    const int64_t n = h->g.n;
    double* outs[7] = {rho, u, v, w, fx, fy, fz};
    for (int k = 0; k < 7; k++)
        if (outs[k]) std::memcpy(outs[k], h->snap_host + k * n, n * 8);
    if (step) *step = h->snap_step;
    return LBM_OK;
}

int lbm_get_force(lbm_handle* h, double* fx, double* fy, double* fz)
{
    return lbm_get_force_component(h, 0, fx, fy, fz);
}

int lbm_get_force_component(lbm_handle* h, int comp, double* fx, double* fy, double* fz)
{
    if (!h) return fail(LBM_EINVAL, "lbm_get_force: NULL handle");
    int rc = check_comp(h, comp);
    if (rc) return rc;
    if (!h->initialised) return fail(LBM_ESTATE, "lbm_get_force before initialisation");
    const int64_t n = h->g.n;
    const int base = 5 * comps(h) + 3;   // force_to_device's staging
    rc = ensure_tmp(h, (base + 3) * n * 8);
    if (rc) return rc;
    double* out = h->tmp + base * n;
    rc = force_to_device(h, comp, out);
    if (rc) return rc;
    double* outs[3] = {fx, fy, fz};
    return arrays_to_host(h, out, outs, 3);
}

int lbm_get_f(lbm_handle* h, double* f) { return lbm_get_f_component(h, 0, f); }

int lbm_get_f_component(lbm_handle* h, int comp, double* f)
{
    if (!h || !f) return fail(LBM_EINVAL, "lbm_get_f: NULL argument");
    int rc = check_comp(h, comp);
    if (rc) return rc;
    if (!h->initialised) return fail(LBM_ESTATE, "lbm_get_f before initialisation");
    const int64_t n = h->g.n * h->d.q;
    rc = ensure_tmp(h, n * 8);
    if (rc) return rc;
    if (!is_lowp(h))
        HIP_OK(lbm::fp64_get_f(h->a, h->g, pops64(h, h->cur, comp), h->kind[h->cur], h->tmp, h->stream));
    else
        HIP_OK(lbm::lowp_get_f(h->d.precision, h->a, h->g, h->d.rho_shift, h->buf[h->cur],
                               h->kind[h->cur], h->rref[h->cur], h->tmp, h->stream));
    HIP_OK(hipMemcpyAsync(f, h->tmp, n * 8, hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    return LBM_OK;
}

int lbm_set_f(lbm_handle* h, const double* f) { return lbm_set_f_component(h, 0, f); }

int lbm_set_f_component(lbm_handle* h, int comp, const double* f)
{
    if (!h || !f) return fail(LBM_EINVAL, "lbm_set_f: NULL argument");
    int rc = check_comp(h, comp);
    if (rc) return rc;
    const int64_t n = h->g.n * h->d.q;
    rc = ensure_tmp(h, 2 * n * 8);
    if (rc) return rc;
    if (comps(h) == 2 && h->initialised && (h->cur != 0 || h->kind[h->cur] != lbm::kHoldsF)) {
        // the other component goes into buffer 0 as f too: its pulled values
        // (post-collision values read upstream) through the staging buffer
        const int other = 1 - comp;
        HIP_OK(lbm::fp64_get_f(h->a, h->g, pops64(h, h->cur, other), h->kind[h->cur], h->tmp + n, h->stream));
        HIP_OK(lbm::fp64_set_f(h->a, h->g, h->tmp + n, pops64(h, 0, other), h->stream));
    }
    HIP_OK(hipMemcpyAsync(h->tmp, f, n * 8, hipMemcpyHostToDevice, h->stream));
    if (!is_lowp(h))
        HIP_OK(lbm::fp64_set_f(h->a, h->g, h->tmp, pops64(h, 0, comp), h->stream));
    else
        HIP_OK(lbm::lowp_set_f(h->d.precision, h->a, h->g, h->d.rho_shift, h->tmp, h->buf[0],
                               h->rref[0], h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    h->cur = 0;
    h->kind[0] = lbm::kHoldsF;
    h->steps = 0;
    h->initialised = true;
    h->first_step_stored = false;
    h->prev_valid = false;
    h->ref_is_rho0 = false;
    return LBM_OK;
}

int lbm_conservation(lbm_handle* h, double out[4]) { return lbm_conservation_component(h, 0, out); }

int lbm_conservation_component(lbm_handle* h, int comp, double out[4])
{
    if (!h || !out) return fail(LBM_EINVAL, "lbm_conservation: NULL argument");
    int rc = check_comp(h, comp);
    if (rc) return rc;
    if (!h->initialised) return fail(LBM_ESTATE, "lbm_conservation before initialisation");
    int nb = 0;
    const int maxb = lbm::fp64_conservation_blocks(h->g);
    rc = ensure_tmp(h, int64_t(maxb) * 4 * 8);
    if (rc) return rc;
    if (!is_lowp(h))
        HIP_OK(lbm::fp64_conservation(h->a, h->g, pops64(h, h->cur, comp), h->kind[h->cur], h->tmp, &nb,
                                      h->stream));
    else
        HIP_OK(lbm::lowp_conservation(h->d.precision, h->a, h->g, h->d.rho_shift,
                                      h->buf[h->cur], h->kind[h->cur], h->rref[h->cur], h->tmp,
                                      &nb, h->stream));
    std::vector<double> part(static_cast<size_t>(nb) * 4);
    HIP_OK(hipMemcpyAsync(part.data(), h->tmp, part.size() * 8, hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    for (int k = 0; k < 4; k++) out[k] = 0.0;
    for (int b = 0; b < nb; b++)
        for (int k = 0; k < 4; k++) out[k] += part[static_cast<size_t>(b) * 4 + k];
    return LBM_OK;
}

int64_t lbm_steps_done(const lbm_handle* h) { return h ? h->steps : -1; }

int lbm_capsule_attach(lbm_handle* h, const lbm_capsule_desc* d)
{
    if (!h || !d) return fail(LBM_EINVAL, "lbm_capsule_attach: NULL argument");
    if (h->has_capsule) return fail(LBM_ESTATE, "a capsule is already attached");
    if (h->initialised) return fail(LBM_ESTATE, "lbm_capsule_attach after initialisation");
    if (h->d.q != 9 || h->d.precision != LBM_FP64 || comps(h) != 1 || h->d.order != 2)
        return fail(LBM_EINVAL, "the capsule model needs a one-component D2Q9 FP64 lattice in step order 2");
    if (h->d.kernel == LBM_KERNEL_FUSED) return fail(LBM_EINVAL, "the capsule model runs the two-pass kernels");
    if (d->nnodes < 1) return fail(LBM_EINVAL, "nnodes must be >= 1 (got %d)", d->nnodes);
    if (!std::isfinite(d->kstiff) || !std::isfinite(d->dt) || !std::isfinite(d->body_fx) ||
        !std::isfinite(d->body_fy))
        return fail(LBM_EINVAL, "capsule parameters must be finite");
    const int64_t n = h->g.n, nn = d->nnodes, nx = h->d.nx;
    const int64_t count = 8 * nn + 2 * n + 4 * nx;
    double* m = nullptr;
    HIP_OK(lbm::dev_malloc(&m, count * 8));
    HIP_OK(hipMemsetAsync(m, 0, count * 8, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    lbm::Capsule& c = h->cap;
    c.x = m; c.y = m + nn; c.x0 = m + 2 * nn; c.y0 = m + 3 * nn;
    c.vx = m + 4 * nn; c.vy = m + 5 * nn; c.nfx = m + 6 * nn; c.nfy = m + 7 * nn;
    c.fx = m + 8 * nn; c.fy = c.fx + n;
    c.gfx = c.fy + n; c.gfy = c.gfx + 2 * nx;
    c.nnodes = d->nnodes;
    c.kstiff = d->kstiff;
    c.dt = d->dt;
    c.body_fx = d->body_fx;
    c.body_fy = d->body_fy;
    h->capsule_mem = m;
    h->has_capsule = true;
    h->capsule_nodes_set = false;
    h->kernel = LBM_KERNEL_TWO_PASS;
    h->block = -1;
    h->bytes += count * 8;
    return LBM_OK;
}

int lbm_capsule_set_nodes(lbm_handle* h, const double* x, const double* y)
{
    if (!h || !x || !y) return fail(LBM_EINVAL, "lbm_capsule_set_nodes: NULL argument");
    if (!h->has_capsule) return fail(LBM_ESTATE, "no capsule attached");
    const int nn = h->cap.nnodes;
    for (int k = 0; k < nn; k++)
        if (!(y[k] >= -1.0 && y[k] < static_cast<double>(h->d.ny)) || !std::isfinite(x[k]))
            return fail(LBM_EINVAL, "node %d at (%g, %g): x must be finite and -1 <= y < ny", k, x[k], y[k]);
    HIP_OK(hipStreamSynchronize(h->stream));
    HIP_OK(hipMemcpy(h->cap.x, x, nn * 8, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(h->cap.y, y, nn * 8, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(h->cap.x0, x, nn * 8, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(h->cap.y0, y, nn * 8, hipMemcpyHostToDevice));
    h->capsule_nodes_set = true;
    return LBM_OK;
}

int lbm_capsule_get_nodes(lbm_handle* h, double* x, double* y, double* vx, double* vy, double* fx, double* fy)
{
    if (!h) return fail(LBM_EINVAL, "lbm_capsule_get_nodes: NULL handle");
    if (!h->has_capsule) return fail(LBM_ESTATE, "no capsule attached");
    const int nn = h->cap.nnodes;
    double* outs[6] = {x, y, vx, vy, fx, fy};
    const double* srcs[6] = {h->cap.x, h->cap.y, h->cap.vx, h->cap.vy, h->cap.nfx, h->cap.nfy};
    for (int k = 0; k < 6; k++)
        if (outs[k]) HIP_OK(hipMemcpyAsync(outs[k], srcs[k], nn * 8, hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    return LBM_OK;
}

int lbm_capsule_nodes(const lbm_handle* h) { return (h && h->has_capsule) ? h->cap.nnodes : 0; }

int lbm_set_tuning(lbm_handle* h, const lbm_tuning* t)
{
    if (!h || !t) return fail(LBM_EINVAL, "lbm_set_tuning: NULL argument");
    if (h->kernel != LBM_KERNEL_FUSED)
        return fail(LBM_EINVAL, "lbm_set_tuning: handle does not use the fused kernel");
    const bool fp64 = h->d.precision == LBM_FP64;
    if (lbm::fused::is_block_shape(t->shape)) {
        // a temporal block: K steps per launch; single steps keep the row shape
        if (!fp64 || comps(h) != 1 || h->has_capsule ||
            !lbm::fused::block_supported(h->g, h->d.q, h->d.walls != 0, t->shape))
            return fail(LBM_EINVAL, "temporal block %d needs a periodic one-component D2Q9 FP64 lattice "
                                    "whose sides it divides", t->shape);
        h->block = t->shape;
        return LBM_OK;
    }
    // shape -1: the defaults, a temporal block included where one applies
    h->block = (t->shape < 0 && fp64 && comps(h) == 1 && !h->has_capsule)
                   ? lbm::fused::default_block(h->g, h->d.q, h->d.walls != 0)
                   : -1;
    const int shape = t->shape >= 0 ? t->shape
                      : fp64    ? lbm::fp64_default_shape(h->g, h->d.q)
                                : lbm::lowp_default_shape(h->d.precision, h->a, h->g);
    if (fp64 ? !lbm::fp64_shape_ok(h->g, h->d.q, shape)
             : !lbm::lowp_shape_ok(h->d.precision, h->a, h->g, shape))
        return fail(LBM_EINVAL, "tile shape %d does not fit the %d x %d cross-section", shape, h->d.ny,
                    h->d.nz);
    if (t->chunk < 1) return fail(LBM_EINVAL, "chunk must be >= 1");
    if (t->order < 0 || t->order > 3) return fail(LBM_EINVAL, "order must be 0, 1, 2 or 3");
    h->tune.shape = shape;
    h->tune.chunk = t->chunk;
    h->tune.order = t->order;
    h->tune.nts = t->nt_store ? 1 : 0;
    return LBM_OK;
}

int lbm_get_tuning(const lbm_handle* h, lbm_tuning* t)
{
    if (!h || !t) return fail(LBM_EINVAL, "lbm_get_tuning: NULL argument");
    t->shape = h->block >= 0 ? h->block : h->tune.shape;
    t->chunk = h->tune.chunk;
    t->order = h->tune.order;
    t->nt_store = h->tune.nts;
    return LBM_OK;
}

int lbm_set_stream(lbm_handle* h, void* stream)
{
    if (!h) return fail(LBM_EINVAL, "lbm_set_stream: NULL handle");
    HIP_OK(hipStreamSynchronize(h->stream));
    h->stream = stream ? static_cast<hipStream_t>(stream) : h->own;
    return LBM_OK;
}

void* lbm_get_stream(const lbm_handle* h) { return h ? static_cast<void*>(h->stream) : nullptr; }

int lbm_synchronize(lbm_handle* h)
{
    if (!h) return fail(LBM_EINVAL, "lbm_synchronize: NULL handle");
    HIP_OK(hipStreamSynchronize(h->stream));
    return LBM_OK;
}

int64_t lbm_device_bytes(const lbm_handle* h) { return h ? h->bytes : 0; }

int lbm_kernel_in_use(const lbm_handle* h) { return h ? h->kernel : -1; }

int64_t lbm_testing_alloc_count(void) { return g_alloc_count.load(); }

void lbm_testing_fail_alloc(int64_t after) { g_fail_after.store(after < 0 ? -1 : after); }

int lbm_fill_noise(uint32_t seed, int64_t n, double mean, double amp, double* out)
{
    if (!out || n < 0) return fail(LBM_EINVAL, "lbm_fill_noise: bad argument");
    std::mt19937 gen(seed);
    for (int64_t c = 0; c < n; c++) {
        const double x = static_cast<double>(gen()) / 4294967295.0;
        out[c] = mean + amp * (x - 0.5);
    }
    return LBM_OK;
}

int lbm_fill_droplet(int nx, int ny, int nz, double R, double W, double rl, double rv, double* out)
{
    if (!out || nx < 1 || ny < 1 || nz < 1 || !(W > 0.0))
        return fail(LBM_EINVAL, "lbm_fill_droplet: bad argument");
    const double cx = 0.5 * (nx - 1), cy = 0.5 * (ny - 1), cz = 0.5 * (nz - 1);
    int64_t c = 0;
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) {
                const double dx = i - cx, dy = j - cy, dz = k - cz;
                const double r = std::sqrt(dx * dx + dy * dy + dz * dz);
                out[c++] = 0.5 * (rl + rv) - 0.5 * (rl - rv) * std::tanh(2.0 * (r - R) / W);
            }
    return LBM_OK;
}

}  // extern "C"
