// lbm_batch.hip -- kernels of the batched engine (lbm_batch.h).
//
// Every job holds post-collision populations as stored values d relative to
// a reference density R of the node that will read them, scaled by S:
// f = d/S + w*R.  Constant shift: R = rs (rs = 0 for FP32 storage), S = 1.
// Node shift (FP16N / BF16N): R is the reading node's density after the
// previous step, an FP32 field double-buffered with the populations (job.R),
// and FP16 values are stored times S = 2^10 (the scheme of the
// single-lattice kernels, lbm_fused_lp.h).  Each step is the single-lattice
// FP32-compute step: pull f(x) = f*(x - e) with periodic wrap, rho = R +
// sum d/S, j = sum d e/S, psi = 1 - exp(-rho), Shan-Chen force
// (scmp3D.cpp:141-179), BGK collision with velocity shift
// (LBfluid3D.cpp:259-276).  A job's first step follows the reference's step
// order for its stencil: D2Q9 jobs collide with the stored rho0 and u = 0
// (LBfluid2D order, LBfluid2D.cpp:235-249 after setFtoFeq), D3Q19 jobs with
// the moments of f0 (scmp3D order, scmp3D.cpp:90-119).
//
// Launches are persistent: a fixed grid whose workgroups take items from a
// ticket counter until the table is exhausted, so uneven jobs balance.  The
// counter used by launch parity p is tickets[p]; each launch clears the other
// one, which the next launch (parity p^1) then starts from.
//
// The row-marching and two-pass kernels evaluate the same expressions in the
// same order (shared helpers below; the file is compiled without FP
// contraction), so a job's bits do not depend on which of them runs.
#include <algorithm>
#include <type_traits>

#include "../../include/mesobasic_lbm.h"
#include "lbm_batch.h"
#include "lbm_fused_lp.h"
#include "lbm_internal.h"
#include "lbm_stencil.h"
#include "lbm_storage.h"

namespace lbm {
namespace batch {
namespace {

__device__ __forceinline__ float psi_of(float rho) { return 1.0f - __expf(-rho); }

// Storage scale of the stored values.
template <typename SP, bool NS>
constexpr float kScale = (NS && std::is_same_v<SP, StF16>) ? 1024.0f : 1.0f;

// Reference density of buffer b at node c.
template <bool NS> __device__ __forceinline__ float ref(const Job& j, int b, int64_t c)
{
    if constexpr (NS) return j.R[b][c];
    else return j.rs;
}

// Density of pulled values: R + (sum d)/S, d summed in population order.
template <int Q, typename SP, bool NS>
__device__ __forceinline__ float pulled_rho(const float (&d)[Q], float R)
{
    float r = 0.f;
    static_for<Q>([&](auto I) { r += d[decltype(I)::value]; });
    return fmaf(r, 1.0f / kScale<SP, NS>, R);
}

// The collision of one node, shared by every batch kernel (D2Q9 row march,
// temporal blocks, two passes of either stencil), so that a job's bits do
// not depend on which of them runs it.  The kernels are VALU-bound (the
// D2Q9 row kernel's vector ALU was busy 62% of its time on MI355X), so the
// terms are arranged for few instructions: one approximate reciprocal of
// rho (1 ulp) instead of two IEEE divides, fused multiply-adds throughout,
// and per population
//   stored_n = d_n omt + (S w_n rho omega) pm1_n + S w_n (omega (rho - Rr) + Rr - Rw)
// with omega = 1/tau, omt = 1 - omega and pm1_n = poly_n - 1 = 3 ev +
// 4.5 ev^2 - 1.5 |u_eq|^2.  That is S (f*_n - w_n Rw) for f*_n = f_n - (f_n -
// feq_n)/tau, f_n = d_n/S + w_n Rr: d_n is the pulled stored value (relative
// to the reading node's Rr), Rw the reference density of the destination
// (Rr = Rw = rs for the constant shift).

// Force sum sum_n w_n e_n psi(x + e_n) over n >= 1 in population order;
// psi_at(I) gives psi at the neighbour of population I.
template <int Q, typename PF>
__device__ __forceinline__ void force_sum(PF psi_at, float& gx, float& gy, float& gz)
{
    using S = St<Q>;
    gx = 0.f; gy = 0.f; gz = 0.f;
    static_for<Q>([&](auto I) {
        constexpr int n = decltype(I)::value;
        if constexpr (n > 0) {
            const float p = psi_at(I);
            constexpr float w = static_cast<float>(S::w[n]);
            if constexpr (S::ex[n] != 0) gx = fmaf(S::ex[n] * w, p, gx);
            if constexpr (S::ey[n] != 0) gy = fmaf(S::ey[n] * w, p, gy);
            if constexpr (S::ez[n] != 0) gz = fmaf(S::ez[n] * w, p, gz);
        }
    });
}

// Per-node terms: equilibrium velocity, 1.5 |u_eq|^2, S rho omega and the
// constant part omega (rho - Rr) (+ Rr with the node shift).
struct BColl {
    float ue, ve, we, usq15, rom, ct, omt;
};

// rho: the node's density (rho0 at a first step in the LBfluid2D order,
// which takes u = 0: at_rest); psi0: psi at the node; g: force_sum.
template <int Q, typename SP, bool NS>
__device__ __forceinline__ BColl batch_prep(const float (&d)[Q], float rho, bool at_rest, float psi0,
                                            float gx, float gy, float gz, float Rr, float tau, float G,
                                            float om, float omt)
{
    using S = St<Q>;
    float jx = 0.f, jy = 0.f, jz = 0.f;
    static_for<Q>([&](auto I) {
        constexpr int n = decltype(I)::value;
        acc_e<S::ex[n]>(jx, d[n]);
        acc_e<S::ey[n]>(jy, d[n]);
        acc_e<S::ez[n]>(jz, d[n]);
    });
    const float ir = __builtin_amdgcn_rcpf(rho);
    const float js = at_rest ? 0.f : ir * (1.0f / kScale<SP, NS>);   // j is summed from scaled values
    const float kf = (-G * psi0) * (tau * ir);
    BColl k;
    k.ue = fmaf(kf, gx, jx * js);
    k.ve = fmaf(kf, gy, jy * js);
    k.we = S::D == 3 ? fmaf(kf, gz, jz * js) : 0.f;
    k.usq15 = 1.5f * fmaf(k.ue, k.ue, fmaf(k.ve, k.ve, k.we * k.we));
    k.rom = rho * om;
    k.ct = om * (rho - Rr);
    k.omt = omt;
    return k;
}

// Stored value of population n after the collision (Rw: reference density
// of the destination node x + e_n; unused for the constant shift).
template <int Q, typename SP, bool NS, int n>
__device__ __forceinline__ float batch_out(const BColl& k, float dn, float Rr, float Rw)
{
    using S = St<Q>;
    constexpr float sw = kScale<SP, NS> * static_cast<float>(S::w[n]);
    const float ev = edot<S::ex[n], S::ey[n], S::ez[n]>(k.ue, k.ve, k.we);
    const float pm1 = fmaf(fmaf(4.5f, ev, 3.0f), ev, -k.usq15);
    float c = k.ct;
    if constexpr (NS) c += Rr - Rw;
    return fmaf(dn, k.omt, fmaf(sw * k.rom, pm1, sw * c));
}

__device__ __forceinline__ int wrapi(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

// Takes the next item of this launch for the whole workgroup; -1 when done.
__device__ __forceinline__ int next_item(const Tables& t, int parity, int* slot)
{
    __syncthreads();
    if (threadIdx.x == 0) *slot = static_cast<int>(atomicAdd(&t.tickets[parity], 1u));
    __syncthreads();
    const int it = *slot;
    return it < *t.nitems ? it : -1;
}

__device__ __forceinline__ void reset_other(const Tables& t, int parity)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) t.tickets[parity ^ 1] = 0u;
}

// ---------------------------------------------------------------------------
// D2Q9: one workgroup marches a job along x with one thread per node of a
// whole y row (periodic in y within the row, so there is no halo ring); the
// psi (and, with the node shift, rho) of rows x-1 .. x+2 sits in four LDS
// slots, and the row x+2 pulls are in flight during the collision of row x.

template <typename SP>
__device__ __forceinline__ void pull_row9(const typename SP::T* __restrict__ P, int64_t ps, int nx,
                                          int ny, int x, int y, float (&d)[9])
{
    using S = St<9>;
    const int xm = x == 0 ? nx - 1 : x - 1, xp = x == nx - 1 ? 0 : x + 1;
    const int ym = y == 0 ? ny - 1 : y - 1, yp = y == ny - 1 ? 0 : y + 1;
    const int xs[3] = {xm, x, xp};
    const int ys[3] = {ym, y, yp};
    // 32-bit node offsets from the plane bases (uniform across the
    // workgroup: one job per item), so each gather takes one VGPR of address
    // and no 64-bit arithmetic; a job holds fewer than 2^30 nodes
    static_for<9>([&](auto I) {
        constexpr int n = decltype(I)::value;
        const uint32_t src = static_cast<uint32_t>(xs[1 - S::ex[n]]) * static_cast<uint32_t>(ny) +
                             static_cast<uint32_t>(ys[1 - S::ey[n]]);
        d[n] = SP::d(*fused::elem(P + n * ps, src));
    });
}

template <typename SP, bool NS, int T>
__global__ void __launch_bounds__(T) k_d2q9_rows(Tables t, int step)
{
    const int parity = step & 1;
    using S = St<9>;
    __shared__ float psi_lds[4 * T];
    __shared__ float rho_lds[NS ? 4 * T : 1];
    __shared__ int slot_s;
    reset_other(t, parity);
    const int y = threadIdx.x;
    for (int it = next_item(t, parity, &slot_s); it >= 0; it = next_item(t, parity, &slot_s)) {
        const Item item = t.items[it];
        Job& job = t.jobs[item.job];
        if (job.remaining <= step) continue;              // uniform across the workgroup
        const int p = parity ^ job.phase;
        const auto* Pin = static_cast<const typename SP::T*>(job.P[p]);
        auto* Pout = static_cast<typename SP::T*>(job.P[p ^ 1]);
        float* Rout = NS ? job.R[p ^ 1] : nullptr;
        const int nx = job.nx, ny = job.ny;
        const int64_t ps = job.ps;
        const float tau = job.tau, G = job.G, om = 1.0f / tau, omt = 1.0f - om;
        // constant shift: one register for the whole item (the row values
        // below then all name it; re-reading job.rs per row, which the
        // stores may alias, cost 5 VGPRs and a wave of occupancy)
        const float rs = job.rs;
        const bool first = job.first != 0 && step == 0;
        const float* rho0 = job.rho0;
        const bool active = y < ny;
        const int ym = y == 0 ? ny - 1 : y - 1, yp = y == ny - 1 ? 0 : y + 1;
        const int x0 = item.x0, x1 = item.x1;
        auto slot = [&](int x) { return ((x - x0 + 1) & 3) * T; };
        auto xw = [&](int x) { return x >= nx ? x - nx : (x < 0 ? x + nx : x); };
        // window entries of pulled row xu (unwrapped): psi, and rho with the
        // node shift (also written to the output reference field for the
        // rows of this item)
        auto put = [&](int xu, const float (&d)[9], float R) {
            const int xr = xw(xu);
            const uint32_t c = static_cast<uint32_t>(xr) * static_cast<uint32_t>(ny) + static_cast<uint32_t>(y);
            const float rp = pulled_rho<9, SP, NS>(d, R);
            psi_lds[slot(xu) + y] = psi_of((first && rho0) ? *fused::elem(rho0, c) : rp);
            if constexpr (NS) {
                rho_lds[slot(xu) + y] = rp;
                if (xu >= x0 && xu < x1) *fused::elem(Rout, c) = rp;
            }
        };
        auto pullR = [&](int xr, float (&d)[9]) {
            pull_row9<SP>(Pin, ps, nx, ny, xr, y, d);
            if constexpr (NS)
                return *fused::elem(static_cast<const float*>(job.R[p]),
                                    static_cast<uint32_t>(xr) * static_cast<uint32_t>(ny) + static_cast<uint32_t>(y));
            else return rs;
        };
        // prologue: the three rows' pulls are all issued before the first
        // window entry is made, so they are one round of loads, not three
        float fa[9], fb[9], fc[9];
        float Ra = 0.f, Rb = 0.f, Rc = 0.f;
        if (active) {
            Rc = pullR(xw(x0 - 1), fc);
            Ra = pullR(x0, fa);
            Rb = pullR(xw(x0 + 1), fb);
            put(x0 - 1, fc, Rc);
            put(x0, fa, Ra);
            put(x0 + 1, fb, Rb);
        }
        __syncthreads();
        auto step = [&](int x, float (&fcur)[9], float Rcur, float (&fnn)[9], float& Rnn) {
            const int x2 = xw(x + 2);
            if (active) {
                Rnn = pullR(x2, fnn);
                const float* pm = psi_lds + slot(x - 1);
                const float* p0 = psi_lds + slot(x);
                const float* pp = psi_lds + slot(x + 1);
                const bool at_rest = first && rho0;
                const float rho = at_rest ? *fused::elem(rho0, static_cast<uint32_t>(x) * static_cast<uint32_t>(ny) +
                                                                   static_cast<uint32_t>(y))
                                          : pulled_rho<9, SP, NS>(fcur, Rcur);
                const int yy[3] = {ym, y, yp};
                float gx, gy, gz;
                force_sum<9>([&](auto I) {
                    constexpr int n = decltype(I)::value;
                    const float* pl = S::ex[n] < 0 ? pm : (S::ex[n] > 0 ? pp : p0);
                    return pl[yy[1 + S::ey[n]]];
                }, gx, gy, gz);
                const BColl k = batch_prep<9, SP, NS>(fcur, rho, at_rest, p0[y], gx, gy, 0.f, Rcur, tau, G,
                                                      om, omt);
                const uint32_t c = static_cast<uint32_t>(x) * static_cast<uint32_t>(ny) + static_cast<uint32_t>(y);
                static_for<9>([&](auto I) {
                    constexpr int n = decltype(I)::value;
                    float Rw = 0.f;
                    if constexpr (NS) {
                        const float* rl = rho_lds + slot(x + S::ex[n]);
                        Rw = rl[yy[1 + S::ey[n]]];
                    }
                    const float v = batch_out<9, SP, NS, n>(k, fcur[n], Rcur, Rw);
                    SP::put_nt(fused::elem(Pout + n * ps, c), SP::of_d(v));
                });
                put(x + 2, fnn, Rnn);
            }
            __syncthreads();
        };
        int x = x0;
        for (; x + 2 < x1; x += 3) {
            step(x, fa, Ra, fc, Rc);
            step(x + 1, fb, Rb, fa, Ra);
            step(x + 2, fc, Rc, fb, Rb);
        }
        if (x < x1) step(x, fa, Ra, fc, Rc);
        if (x + 1 < x1) step(x + 1, fb, Rb, fa, Ra);
    }
}


// ---------------------------------------------------------------------------
// D2Q9 temporal blocks (constant shift): an item is a T x T tile of one job.
// The stored values of the tile and a halo of H = 2K nodes (wrapped
// periodically) go to LDS once; K steps run there (the region whose values
// are valid shrinks by 2 per step: f is pulled from radius 1, psi of the
// neighbours from radius 2), each step decoding the pulled stored values,
// forming rho and psi (barrier), then the force, collision and the stored
// value of the result, written back in place (all pulls of the step are
// done at the barrier; a second barrier ends the step).  The arithmetic and
// the rounding to the stored type are those of k_d2q9_rows, so K steps here
// give the bits of K single steps.  Tiles past the lattice end wrap; their
// extra nodes repeat nodes of the first tiles and are not stored.

template <typename SP, int T, int K>
__global__ void __launch_bounds__(1024) k_d2q9_block(Tables t, int parity)
{
    using S = St<9>;
    using ST = typename SP::T;
    constexpr int H = 2 * K, W = T + 2 * H, W2 = W * W;
    constexpr int NPT = (W2 + 1023) / 1024;   // region nodes per thread
    __shared__ __align__(16) unsigned char fs_raw[9 * W2 * sizeof(ST)];
    __shared__ float pw[W2];
    __shared__ float r0[W2];
    __shared__ int slot_s;
    ST* fs = reinterpret_cast<ST*>(fs_raw);
    reset_other(t, parity);
    // The stored values of the next item are loaded into registers while
    // the current one computes, so a tile's load latency is hidden behind
    // the previous tile's steps.
    ST pre[NPT][9];
    float pre0[NPT];
    // region coordinates of this thread's slots (fixed for the launch)
    int si[NPT], sj[NPT];
    for (int q = 0; q < NPT; q++) {
        const int l = threadIdx.x + q * 1024;
        si[q] = l / W;
        sj[q] = l - si[q] * W;
    }
    auto load = [&](int itn) {
        const Item item = t.items[itn];
        const Job& job = t.jobs[item.job];
        if (job.remaining <= 0) return;
        const auto* Pin = static_cast<const ST*>(job.P[job.done & 1]);
        const int nx = job.nx, ny = job.ny;
        const bool first = job.first != 0 && job.rho0 != nullptr;
        for (int q = 0; q < NPT; q++) {
            const int l = threadIdx.x + q * 1024;
            const int i = si[q], j = sj[q];
            int gx = (item.x0 - H + i) % nx, gy = (item.y0 - H + j) % ny;
            gx += gx < 0 ? nx : 0;
            gy += gy < 0 ? ny : 0;
            if (l < W2) {
                const int64_t c = static_cast<int64_t>(gx) * ny + gy;
                static_for<9>([&](auto I) {
                    constexpr int n = decltype(I)::value;
                    pre[q][n] = Pin[n * job.ps + c];
                });
                if (first) pre0[q] = job.rho0[c];
            }
        }
    };
    int it = next_item(t, parity, &slot_s);
    if (it >= 0) load(it);
    while (it >= 0) {
        const Item item = t.items[it];
        const Job& job = t.jobs[item.job];
        const bool live = job.remaining > 0;            // uniform across the workgroup
        const int k = job.remaining > K ? K : 1;
        // K steps while more than K are left, then single steps: the last
        // step before a stop is a single one, so the other buffer then
        // holds f(N-1), as lbm_batch_read expects
        auto* Pout = static_cast<ST*>(job.P[(job.done & 1) ^ 1]);
        const int nx = job.nx, ny = job.ny;
        const int64_t ps = job.ps;
        const float tau = job.tau, G = job.G, om = 1.0f / tau, omt = 1.0f - om, rs = job.rs;
        const bool first = job.first != 0 && job.rho0 != nullptr;
        if (live) {
            for (int q = 0; q < NPT; q++) {
                const int l = threadIdx.x + q * 1024;
                if (l < W2) {
                    static_for<9>([&](auto I) {
                        constexpr int n = decltype(I)::value;
                        fs[n * W2 + l] = pre[q][n];
                    });
                    if (first) r0[l] = pre0[q];
                }
            }
        }
        // next item (its barrier also publishes the LDS writes above), its
        // loads in flight during this item's steps
        const int nit = next_item(t, parity, &slot_s);
        if (nit >= 0) load(nit);
        if (live) {
            float d[NPT][9];
            float rp[NPT];
            for (int s = 0; s < k; s++) {
                const int lo = 2 * s + 1, hi = W - 2 * s - 1;
                const bool f0 = first && s == 0;
                for (int q = 0; q < NPT; q++) {
                    const int l = threadIdx.x + q * 1024;
                    const int i = si[q], j = sj[q];
                    if (l < W2 && i >= lo && i < hi && j >= lo && j < hi) {
                        static_for<9>([&](auto I) {
                            constexpr int n = decltype(I)::value;
                            d[q][n] = SP::d(fs[n * W2 + l - (S::ex[n] * W + S::ey[n])]);
                        });
                        rp[q] = pulled_rho<9, SP, false>(d[q], rs);
                        pw[l] = psi_of(f0 ? r0[l] : rp[q]);
                    }
                }
                __syncthreads();
                for (int q = 0; q < NPT; q++) {
                    const int l = threadIdx.x + q * 1024;
                    const int i = si[q], j = sj[q];
                    if (l < W2 && i > lo && i < hi - 1 && j > lo && j < hi - 1) {
                        const float rho = f0 ? r0[l] : rp[q];
                        float gx, gy, gz;
                        force_sum<9>([&](auto I) {
                            constexpr int n = decltype(I)::value;
                            return pw[l + S::ex[n] * W + S::ey[n]];
                        }, gx, gy, gz);
                        const BColl kc = batch_prep<9, SP, false>(d[q], rho, f0, pw[l], gx, gy, 0.f, rs, tau, G,
                                                                  om, omt);
                        static_for<9>([&](auto I) {
                            constexpr int n = decltype(I)::value;
                            fs[n * W2 + l] = SP::of_d(batch_out<9, SP, false, n>(kc, d[q][n], rs, rs));
                        });
                    }
                }
                __syncthreads();
            }
            for (int q = 0; q < NPT; q++) {
                const int l = threadIdx.x + q * 1024;
                const int i = si[q], j = sj[q];
                const int xu = item.x0 - H + i, yu = item.y0 - H + j;
                if (l < W2 && i >= H && i < W - H && j >= H && j < W - H && xu < nx && yu < ny) {
                    const int64_t c = static_cast<int64_t>(xu) * ny + yu;
                    static_for<9>([&](auto I) {
                        constexpr int n = decltype(I)::value;
                        SP::put_nt(Pout + n * ps + c, fs[n * W2 + l]);
                    });
                }
            }
        }
        __syncthreads();   // the stores read fs before the next item's values replace it
        it = nit;
    }
}

// ---------------------------------------------------------------------------
// Two-pass kernels for any job (D2Q9 or D3Q19, any size): items are node
// ranges; pass 1 writes psi (and the node-shift reference), pass 2 collides.

template <int Q>
__device__ __forceinline__ void decompose(int64_t c, const Job& j, int& x, int& y, int& z)
{
    const int64_t t = c / j.nz;
    z = static_cast<int>(c - t * j.nz);
    x = static_cast<int>(t / j.ny);
    y = static_cast<int>(t - static_cast<int64_t>(x) * j.ny);
}

template <int Q, typename SP>
__device__ __forceinline__ void pull_node(const typename SP::T* __restrict__ P, const Job& j, int x,
                                          int y, int z, float (&d)[Q])
{
    using S = St<Q>;
    const int xs[3] = {x == 0 ? j.nx - 1 : x - 1, x, x == j.nx - 1 ? 0 : x + 1};
    const int ys[3] = {y == 0 ? j.ny - 1 : y - 1, y, y == j.ny - 1 ? 0 : y + 1};
    const int zs[3] = {z == 0 ? j.nz - 1 : z - 1, z, z == j.nz - 1 ? 0 : z + 1};
    static_for<Q>([&](auto I) {
        constexpr int n = decltype(I)::value;
        const int64_t src = (static_cast<int64_t>(xs[1 - S::ex[n]]) * j.ny + ys[1 - S::ey[n]]) * j.nz +
                            zs[1 - S::ez[n]];
        d[n] = SP::d(P[n * j.ps + src]);
    });
}

// The two-pass kernels address in 32-bit node indices: a job holds fewer
// than 2^30 nodes (lbm_batch_submit), so byte offsets of the 2- and 4-byte
// values fit 32 bits and every gather takes the SGPR-base + 32-bit VGPR
// offset form (fused::elem) off the job's wave-uniform buffer pointers.
// 64-bit node arithmetic (two 64-bit divides per node, a 64-bit
// multiply-add per gather) held 2 VGPRs per address.
template <int Q>
__device__ __forceinline__ void decompose32(uint32_t c, const Job& j, int& x, int& y, int& z)
{
    const uint32_t ny = static_cast<uint32_t>(j.ny);
    uint32_t t = c;
    z = 0;
    if constexpr (Q == 19) {
        const uint32_t nz = static_cast<uint32_t>(j.nz);
        t = c / nz;
        z = static_cast<int>(c - t * nz);
    }
    x = static_cast<int>(t / ny);
    y = static_cast<int>(t - static_cast<uint32_t>(x) * ny);
}

__device__ __forceinline__ fused::NbrOff neighbour_offsets(const Job& j, int x, int y, int z)
{
    Grid g;
    g.nx = j.nx; g.ny = j.ny; g.nz = j.nz; g.n = j.n; g.ps = j.ps;
    return fused::NbrOff(g, x, y, z);
}

// Pulled stored values of node c: population n from node c - e_n.
template <int Q, typename SP>
__device__ __forceinline__ void pull_node32(const typename SP::T* __restrict__ P, const Job& j,
                                            uint32_t c, const fused::NbrOff& o, float (&d)[Q])
{
    using S = St<Q>;
    static_for<Q>([&](auto I) {
        constexpr int n = decltype(I)::value;
        d[n] = SP::d(*fused::elem(P + n * j.ps, c + o.template to<-S::ex[n], -S::ey[n], -S::ez[n]>()));
    });
}

constexpr int kTP = 256;

template <int Q, typename SP, bool NS>
__global__ void __launch_bounds__(kTP) k_tp_psi(Tables t, int step)
{
    const int parity = step & 1;
    __shared__ int slot_s;
    reset_other(t, parity);
    for (int it = next_item(t, parity, &slot_s); it >= 0; it = next_item(t, parity, &slot_s)) {
        const Item item = t.items[it];
        const Job& job = t.jobs[item.job];
        if (job.remaining <= step) continue;
        const int p = parity ^ job.phase;
        const auto* Pin = static_cast<const typename SP::T*>(job.P[p]);
        const uint32_t c1 = static_cast<uint32_t>(item.c1);
        for (uint32_t c = static_cast<uint32_t>(item.c0) + threadIdx.x; c < c1; c += kTP) {
            const bool from_rho0 = job.first && step == 0 && job.rho0;
            float rp = 0.f;
            if (!from_rho0 || NS) {
                int x, y, z;
                decompose32<Q>(c, job, x, y, z);
                float d[Q];
                pull_node32<Q, SP>(Pin, job, c, neighbour_offsets(job, x, y, z), d);
                float R = job.rs;
                if constexpr (NS) R = *fused::elem(job.R[p], c);
                rp = pulled_rho<Q, SP, NS>(d, R);
            }
            *fused::elem(job.psi, c) = psi_of(from_rho0 ? *fused::elem(job.rho0, c) : rp);
            if constexpr (NS) *fused::elem(job.R[p ^ 1], c) = rp;
        }
    }
}

template <int Q, typename SP, bool NS>
__global__ void __launch_bounds__(kTP) k_tp_collide(Tables t, int step)
{
    const int parity = step & 1;
    using S = St<Q>;
    __shared__ int slot_s;
    reset_other(t, parity);
    for (int it = next_item(t, parity, &slot_s); it >= 0; it = next_item(t, parity, &slot_s)) {
        const Item item = t.items[it];
        const Job& job = t.jobs[item.job];
        if (job.remaining <= step) continue;
        const int p = parity ^ job.phase;
        const auto* Pin = static_cast<const typename SP::T*>(job.P[p]);
        auto* Pout = static_cast<typename SP::T*>(job.P[p ^ 1]);
        const float tau = job.tau, G = job.G, om = 1.0f / tau, omt = 1.0f - om;
        const float* psi = job.psi;
        const float rs = job.rs;
        const uint32_t c1 = static_cast<uint32_t>(item.c1);
        for (uint32_t c = static_cast<uint32_t>(item.c0) + threadIdx.x; c < c1; c += kTP) {
            int x, y, z;
            decompose32<Q>(c, job, x, y, z);
            const fused::NbrOff o = neighbour_offsets(job, x, y, z);
            float d[Q];
            pull_node32<Q, SP>(Pin, job, c, o, d);
            float Rr = rs;
            if constexpr (NS) Rr = *fused::elem(job.R[p], c);
            const bool at_rest = job.first && step == 0 && job.rho0;
            const float rho = at_rest ? *fused::elem(job.rho0, c) : pulled_rho<Q, SP, NS>(d, Rr);
            // node c + e_n
            auto nb = [&](auto N) {
                constexpr int n = decltype(N)::value;
                return c + o.template to<S::ex[n], S::ey[n], S::ez[n]>();
            };
            float gx, gy, gz;
            force_sum<Q>([&](auto I) { return *fused::elem(psi, nb(I)); }, gx, gy, gz);
            const BColl k = batch_prep<Q, SP, NS>(d, rho, at_rest, *fused::elem(psi, c), gx, gy, gz, Rr, tau, G,
                                                  om, omt);
            static_for<Q>([&](auto I) {
                constexpr int n = decltype(I)::value;
                float Rw = 0.f;
                if constexpr (NS) Rw = *fused::elem(job.R[p ^ 1], nb(I));
                const float v = batch_out<Q, SP, NS, n>(k, d[n], Rr, Rw);
                SP::put_nt(fused::elem(Pout + n * job.ps, c), SP::of_d(v));
            });
        }
    }
}

// ---------------------------------------------------------------------------
// D3Q19 march: an item is a kMarchTY x kMarchTZ tile of one job's y-z cross
// section and a chunk of x planes [x0, x1), run by the single-lattice fused
// march (fused::lp_tile, lbm_fused_lp.h): psi (and rho with the node shift)
// of planes x-1 .. x+2 for the tile and its one-node ring in LDS, the pulls
// of plane x+2 in flight during the collision of plane x, every population
// read once per step.  Sides that are not multiples of the tile run as
// partial tiles (wrapped pulls, masked stores).  The stored values follow
// the convention of this engine (d = S (f - w R) relative to the reading
// node's R), which is the single-lattice one.
//
// Its FP32 arithmetic is lp_prep / lp_out, not the expressions of the
// two-pass kernels above, so the two agree to rounding, not bit for bit; a
// D3Q19 job therefore always runs the same kernel, whatever the other jobs
// or the tuner do: which one follows from the job's own shape and the
// batch's mode (use_march() in lbm_batch_capi.cpp).

template <typename SP, bool NS>
// Occupancy target 2 wavefronts per SIMD: held to 128 VGPRs the item loop
// spills 36-108 bytes per thread, and measured slower.
__global__ void __launch_bounds__(kMarchTY * kMarchTZ) __attribute__((amdgpu_waves_per_eu(2)))
    k_d3q19_march(Tables t, int step)
{
    const int parity = step & 1;
    using TL = fused::Tile<kMarchTY, kMarchTZ, false>;
    constexpr int PL = TL::LY * TL::LZ;
    __shared__ float wpsi[4 * PL], wrho[4 * PL];
    __shared__ int slot_s;
    reset_other(t, parity);
    for (int it = next_item(t, parity, &slot_s); it >= 0; it = next_item(t, parity, &slot_s)) {
        const Item item = t.items[it];
        const Job& job = t.jobs[item.job];
        if (job.remaining <= step) continue;              // uniform across the workgroup
        const int p = parity ^ job.phase;
        Grid g;
        g.nx = job.nx;
        g.ny = job.ny;
        g.nz = job.nz;
        g.n = job.n;
        g.ps = job.ps;
        fused::lp_tile<SP, false, NS, false, TL, true, true>(
            static_cast<const typename SP::T*>(job.P[p]), kHoldsPost,
            static_cast<typename SP::T*>(job.P[p ^ 1]), g, job.tau, job.G, job.rs,
            NS ? job.R[p] : nullptr, NS ? job.R[p ^ 1] : nullptr, item.y0, item.z0, item.x0,
            item.x1, wpsi, wrho);
    }
}

// After the single steps of a tick: each job took min(steps, remaining) of
// them (a step launch skips jobs with remaining <= its step index).
__global__ void k_advance(Job* jobs, int njobs, int steps, unsigned* mass_ticket)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *mass_ticket = 0u;
    if (i < njobs && !jobs[i].blk && jobs[i].remaining > 0) {
        const int64_t k = jobs[i].remaining < steps ? jobs[i].remaining : steps;
        jobs[i].remaining -= k;
        jobs[i].done += k;
        jobs[i].first = 0;
    }
}

// After a temporal-block launch: the steps k_d2q9_block ran for each block job.
__global__ void k_advance_blk(Job* jobs, int njobs, unsigned* mass_ticket)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *mass_ticket = 0u;
    if (i < njobs && jobs[i].blk && jobs[i].remaining > 0) {
        const int64_t k = jobs[i].remaining > kBlkK ? kBlkK : 1;
        jobs[i].remaining -= k;
        jobs[i].done += k;
        jobs[i].first = 0;
    }
}

// Per-item FP64 mass of a job's current populations.  Every stored value is
// pulled by exactly one node (periodic wrap is a bijection), so the sum over
// all slots of d/S plus R per node is the total mass of the pulled f; the
// slots are read plane by plane, coalesced.  Tree reduction in the block:
// deterministic for a given item.
constexpr int kMassBlock = 256;

template <typename SP, bool NS>
__global__ void __launch_bounds__(kMassBlock) k_job_mass(Tables t, unsigned* ticket, double* part)
{
    __shared__ int slot_s;
    __shared__ double red[kMassBlock];
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) slot_s = static_cast<int>(atomicAdd(ticket, 1u));
        __syncthreads();
        const int it = slot_s;
        if (it >= *t.nitems) break;
        const Item item = t.items[it];
        const Job& job = t.jobs[item.job];
        const int b = static_cast<int>(job.done & 1);
        const auto* P = static_cast<const typename SP::T*>(job.P[b]);
        const float* R = NS ? job.R[b] : nullptr;
        const int q = job.q;
        const double rs = job.rs;
        double m = 0.0;
        for (int64_t c = item.c0 + threadIdx.x; c < item.c1; c += kMassBlock) {
            double s = 0.0;
            for (int n = 0; n < q; n++) s += static_cast<double>(SP::d(P[n * job.ps + c]));
            m += s * (1.0 / kScale<SP, NS>) + (NS ? static_cast<double>(R[c]) : rs);
        }
        red[threadIdx.x] = m;
        __syncthreads();
        for (int k = kMassBlock / 2; k > 0; k >>= 1) {
            if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
            __syncthreads();
        }
        if (threadIdx.x == 0) part[it] = red[0];
    }
}

// Initial equilibrium: psi(rho0) into job.psi (and R = rho0 for the node
// shift), then the equilibrium of each node, relative to that node's
// reference, scattered to the upstream slots so that its first pull reads
// it back.
template <int Q, bool NS>
__global__ void __launch_bounds__(kTP) k_init_psi(Job j, int buf)
{
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kTP + threadIdx.x;
    if (c >= j.n) return;
    j.psi[c] = psi_of(j.rho0[c]);
    if constexpr (NS) j.R[buf][c] = j.rho0[c];
}

template <int Q, typename SP, bool NS>
__global__ void __launch_bounds__(kTP) k_init_feq(Job j, int buf)
{
    using S = St<Q>;
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kTP + threadIdx.x;
    if (c >= j.n) return;
    auto* P = static_cast<typename SP::T*>(j.P[buf]);
    int x, y, z;
    decompose<Q>(c, j, x, y, z);
    const int xs[3] = {x == 0 ? j.nx - 1 : x - 1, x, x == j.nx - 1 ? 0 : x + 1};
    const int ys[3] = {y == 0 ? j.ny - 1 : y - 1, y, y == j.ny - 1 ? 0 : y + 1};
    const int zs[3] = {z == 0 ? j.nz - 1 : z - 1, z, z == j.nz - 1 ? 0 : z + 1};
    float gx = 0.f, gy = 0.f, gz = 0.f;
    static_for<Q>([&](auto I) {
        constexpr int n = decltype(I)::value;
        if constexpr (n > 0) {
            const int64_t nb = (static_cast<int64_t>(xs[1 + S::ex[n]]) * j.ny + ys[1 + S::ey[n]]) * j.nz +
                               zs[1 + S::ez[n]];
            const float tt = static_cast<float>(S::w[n]) * j.psi[nb];
            acc_e<S::ex[n]>(gx, tt);
            acc_e<S::ey[n]>(gy, tt);
            acc_e<S::ez[n]>(gz, tt);
        }
    });
    const float rho = j.rho0[c];
    const float Rr = ref<NS>(j, buf, c);
    const float k = (-j.G * j.psi[c]) * j.tau / rho;
    const float ue = k * gx, ve = k * gy, we = (S::D == 3) ? k * gz : 0.f;
    const float usq15 = 1.5f * (ue * ue + ve * ve + we * we);
    static_for<Q>([&](auto I) {
        constexpr int n = decltype(I)::value;
        const float ev = edot<S::ex[n], S::ey[n], S::ez[n]>(ue, ve, we);
        const float poly = 1.0f + 3.0f * ev + 4.5f * ev * ev - usq15;
        const float deq = (kScale<SP, NS> * static_cast<float>(S::w[n])) * (rho * poly - Rr);
        // slot read by the pull at node c: upstream node c - e_n
        const int64_t up = (static_cast<int64_t>(xs[1 - S::ex[n]]) * j.ny + ys[1 - S::ey[n]]) * j.nz +
                           zs[1 - S::ez[n]];
        P[n * j.ps + up] = SP::of_d(deq);
    });
}

template <int Q, typename SP, bool NS>
__global__ void __launch_bounds__(kTP) k_job_moments(Job j, int buf, double* __restrict__ rho,
                                                     double* __restrict__ u, double* __restrict__ v,
                                                     double* __restrict__ w)
{
    using S = St<Q>;
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kTP + threadIdx.x;
    if (c >= j.n) return;
    const auto* P = static_cast<const typename SP::T*>(j.P[buf]);
    int x, y, z;
    decompose<Q>(c, j, x, y, z);
    float d[Q];
    pull_node<Q, SP>(P, j, x, y, z, d);
    const double R = ref<NS>(j, buf, c);
    double r = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
    static_for<Q>([&](auto I) {
        constexpr int n = decltype(I)::value;
        const double fv = static_cast<double>(d[n]) / kScale<SP, NS> + S::w[n] * R;
        r += fv;
        acc_e<S::ex[n]>(sx, fv);
        acc_e<S::ey[n]>(sy, fv);
        acc_e<S::ez[n]>(sz, fv);
    });
    rho[c] = r;
    u[c] = sx / r;
    v[c] = sy / r;
    w[c] = (S::D == 3) ? sz / r : 0.0;
}

__global__ void k_widen(const float* __restrict__ a, double* __restrict__ b, int64_t n)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kTP + threadIdx.x;
    if (i < n) b[i] = static_cast<double>(a[i]);
}

__global__ void k_zero(unsigned* p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// Runs f(storage tag, node-shift tag) for a batch precision.
template <typename F> void by_prec(int prec, F&& f)
{
    switch (prec) {
    case LBM_FP32: f(StF32{}, std::false_type{}); break;
    case LBM_FP16S: f(StF16{}, std::false_type{}); break;
    case LBM_BF16S: f(StBF16{}, std::false_type{}); break;
    case LBM_FP16N: f(StF16{}, std::true_type{}); break;
    default: f(StBF16{}, std::true_type{}); break;
    }
}

inline unsigned blocks(int64_t n) { return static_cast<unsigned>((n + kTP - 1) / kTP); }

}  // namespace

hipError_t d2q9_rows_step(int prec, const Tables& t, int step, int threads, int grid,
                          hipStream_t s)
{
    if (threads < 1 || threads > 1024) return hipErrorInvalidValue;   // one thread per row node
    by_prec(prec, [&](auto sp, auto ns) {
        using SP = decltype(sp);
        constexpr bool NS = decltype(ns)::value;
        if (threads <= 64) k_d2q9_rows<SP, NS, 64><<<grid, 64, 0, s>>>(t, step);
        else if (threads <= 128) k_d2q9_rows<SP, NS, 128><<<grid, 128, 0, s>>>(t, step);
        else if (threads <= 256) k_d2q9_rows<SP, NS, 256><<<grid, 256, 0, s>>>(t, step);
        else if (threads <= 512) k_d2q9_rows<SP, NS, 512><<<grid, 512, 0, s>>>(t, step);
        else k_d2q9_rows<SP, NS, 1024><<<grid, 1024, 0, s>>>(t, step);
    });
    return hipGetLastError();
}

hipError_t twopass_psi(int prec, int q, const Tables& t, int step, int grid, hipStream_t s)
{
    by_prec(prec, [&](auto sp, auto ns) {
        using SP = decltype(sp);
        constexpr bool NS = decltype(ns)::value;
        if (q == 9) k_tp_psi<9, SP, NS><<<grid, kTP, 0, s>>>(t, step);
        else k_tp_psi<19, SP, NS><<<grid, kTP, 0, s>>>(t, step);
    });
    return hipGetLastError();
}

hipError_t twopass_collide(int prec, int q, const Tables& t, int step, int grid, hipStream_t s)
{
    by_prec(prec, [&](auto sp, auto ns) {
        using SP = decltype(sp);
        constexpr bool NS = decltype(ns)::value;
        if (q == 9) k_tp_collide<9, SP, NS><<<grid, kTP, 0, s>>>(t, step);
        else k_tp_collide<19, SP, NS><<<grid, kTP, 0, s>>>(t, step);
    });
    return hipGetLastError();
}

hipError_t d3q19_march_step(int prec, const Tables& t, int step, int grid, hipStream_t s)
{
    by_prec(prec, [&](auto sp, auto ns) {
        using SP = decltype(sp);
        constexpr bool NS = decltype(ns)::value;
        k_d3q19_march<SP, NS><<<grid, kMarchTY * kMarchTZ, 0, s>>>(t, step);
    });
    return hipGetLastError();
}

hipError_t advance_counters(Job* jobs, int njobs, int steps, unsigned* mass_ticket, hipStream_t s)
{
    k_advance<<<(std::max(njobs, 1) + 255) / 256, 256, 0, s>>>(jobs, njobs, steps, mass_ticket);
    return hipGetLastError();
}

hipError_t advance_blocks(Job* jobs, int njobs, unsigned* mass_ticket, hipStream_t s)
{
    k_advance_blk<<<(std::max(njobs, 1) + 255) / 256, 256, 0, s>>>(jobs, njobs, mass_ticket);
    return hipGetLastError();
}

hipError_t d2q9_block_step(int prec, const Tables& t, int parity, int grid, hipStream_t s)
{
    if (prec == LBM_FP16N || prec == LBM_BF16N) return hipErrorNotSupported;
    by_prec(prec, [&](auto sp, auto ns) {
        using SP = decltype(sp);
        if constexpr (!decltype(ns)::value)
            k_d2q9_block<SP, kBlkT, kBlkK><<<grid, 1024, 0, s>>>(t, parity);
    });
    return hipGetLastError();
}

hipError_t job_mass(int prec, const Tables& t, unsigned* ticket, double* part, int grid,
                    hipStream_t s)
{
    by_prec(prec, [&](auto sp, auto ns) {
        using SP = decltype(sp);
        constexpr bool NS = decltype(ns)::value;
        k_job_mass<SP, NS><<<grid, kMassBlock, 0, s>>>(t, ticket, part);
    });
    return hipGetLastError();
}

hipError_t init_job(int prec, int q, const Job& j, int buf, hipStream_t s)
{
    by_prec(prec, [&](auto sp, auto ns) {
        using SP = decltype(sp);
        constexpr bool NS = decltype(ns)::value;
        if (q == 9) {
            k_init_psi<9, NS><<<blocks(j.n), kTP, 0, s>>>(j, buf);
            k_init_feq<9, SP, NS><<<blocks(j.n), kTP, 0, s>>>(j, buf);
        } else {
            k_init_psi<19, NS><<<blocks(j.n), kTP, 0, s>>>(j, buf);
            k_init_feq<19, SP, NS><<<blocks(j.n), kTP, 0, s>>>(j, buf);
        }
    });
    return hipGetLastError();
}

hipError_t job_moments(int prec, int q, const Job& j, int buf, double* rho, double* u, double* v,
                       double* w, hipStream_t s)
{
    by_prec(prec, [&](auto sp, auto ns) {
        using SP = decltype(sp);
        constexpr bool NS = decltype(ns)::value;
        if (q == 9) k_job_moments<9, SP, NS><<<blocks(j.n), kTP, 0, s>>>(j, buf, rho, u, v, w);
        else k_job_moments<19, SP, NS><<<blocks(j.n), kTP, 0, s>>>(j, buf, rho, u, v, w);
    });
    return hipGetLastError();
}

hipError_t widen(const float* a, double* b, int64_t n, hipStream_t s)
{
    k_widen<<<blocks(n), kTP, 0, s>>>(a, b, n);
    return hipGetLastError();
}

hipError_t zero_u32(unsigned* p, int n, hipStream_t s)
{
    k_zero<<<(n + 255) / 256, 256, 0, s>>>(p, n);
    return hipGetLastError();
}

}  // namespace batch
}  // namespace lbm
