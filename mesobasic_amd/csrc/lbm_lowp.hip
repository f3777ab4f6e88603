// lbm_lowp.hip -- reduced-precision Shan-Chen step: FP32 arithmetic with
// FP32 storage of f, or FP16 / BF16 storage of shifted populations
// d_i = f_i - w_i*R.  With a constant R (= rho_shift) this is "DDF
// shifting" (Lehmann et al. 2022, "Accuracy and performance of the lattice
// Boltzmann method with 64-bit, 32-bit, and customized 16-bit number
// formats"); the node-shift forms use the reading node's density after the
// previous step instead (lbm_fused_lp.h describes the scheme).  Same step
// structure and formulas as the FP64 path (lbm_fp64.hip, which cites the
// reference); only the number formats differ, so results are compared with
// the FP64 oracle by tolerance, not bit for bit.
//
// This file holds the two-pass kernels (any stencil and lattice), the
// initialisation, read-back and reductions; the fused D3Q19 kernel is
// launched from lbm_lowp_fused_*.hip.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include "../../include/mesobasic_lbm.h"
#include "lbm_lowp.h"
#include "lbm_fused_lp.h"
#include "lbm_stencil.h"
#include "lbm_storage.h"

namespace lbm {
namespace {

constexpr int kBlock = 256;

struct Coord {
    int x, y, z;
};

__device__ __forceinline__ Coord decompose(uint32_t c, const Grid& g)
{
    Coord r;
    const uint32_t nz = static_cast<uint32_t>(g.nz), ny = static_cast<uint32_t>(g.ny);
    const uint32_t t = c / nz;
    r.z = static_cast<int>(c - t * nz);
    r.x = static_cast<int>(t / ny);
    r.y = static_cast<int>(t - static_cast<uint32_t>(r.x) * ny);
    return r;
}

struct Nbr {
    int64_t xy[3][3];
    int z[3];
};

__device__ __forceinline__ Nbr neighbours(const Coord& p, const Grid& g)
{
    Nbr nb;
    const int xs[3] = {p.x == 0 ? g.nx - 1 : p.x - 1, p.x, p.x == g.nx - 1 ? 0 : p.x + 1};
    const int ys[3] = {p.y == 0 ? g.ny - 1 : p.y - 1, p.y, p.y == g.ny - 1 ? 0 : p.y + 1};
    nb.z[0] = p.z == 0 ? g.nz - 1 : p.z - 1;
    nb.z[1] = p.z;
    nb.z[2] = p.z == g.nz - 1 ? 0 : p.z + 1;
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++)
            nb.xy[a][b] = (static_cast<int64_t>(xs[a]) * g.ny + ys[b]) * g.nz;
    return nb;
}

__device__ __forceinline__ int64_t at(const Nbr& nb, int dx, int dy, int dz)
{
    return nb.xy[dx + 1][dy + 1] + nb.z[dz + 1];
}

// Reference density of the stored values: the node-shift field, or the
// constant shift.
template <bool NS> struct Ref {
    const float* r;
    float rs;
    __device__ __forceinline__ float operator()(int64_t c) const
    {
        if constexpr (NS) return r[c];
        else return rs;
    }
};

// Scale of the stored values (FP16 node shift: 2^10, see lbm_fused_lp.h).
template <typename SP, bool NS> constexpr float kScale = fused::kStoreScale<SP, NS>;

// f at node c (coordinates p) from the stored values: gathered from the
// upstream neighbours (kHoldsPost) or in place (kHoldsF), plus w_n*R(c).
// Bounce-back slots were written by c itself for its wall-side neighbour,
// so they carry that neighbour's reference density.
template <int Q, bool WALLS, typename SP, bool NS>
__device__ __forceinline__ void load_f(const typename SP::T* __restrict__ P, const Grid& g, int kind,
                                       const Coord& p, const Nbr& nb, int64_t c, const Ref<NS>& R,
                                       float (&f)[Q])
{
    using S = St<Q>;
    const float Rc = R(c);
    if (kind == kHoldsF) {
        static_for<Q>([&](auto I) {
            constexpr int n = decltype(I)::value;
            f[n] = fmaf(static_cast<float>(S::w[n]), Rc, SP::d(P[n * g.ps + c]) * (1.0f / kScale<SP, NS>));
        });
        return;
    }
    static_for<Q>([&](auto I) {
        constexpr int n = decltype(I)::value;
        f[n] = fmaf(static_cast<float>(S::w[n]), Rc,
                    SP::d(P[n * g.ps + at(nb, -S::ex[n], -S::ey[n], -S::ez[n])]) *
                        (1.0f / kScale<SP, NS>));
    });
    if constexpr (WALLS) {
        const int last = (S::D == 3) ? p.z : p.y;
        const int nl = (S::D == 3) ? g.nz : g.ny;
        auto bounce = [&](auto M, auto from_dn) {
            constexpr int m = decltype(M)::value;
            constexpr bool DN = decltype(from_dn)::value;
            constexpr int src = DN ? S::wall_dn[m] : S::wall_up[m];
            constexpr int dst = DN ? S::wall_up[m] : S::wall_dn[m];
            const float Rn = R(at(nb, S::ex[src], S::ey[src], S::ez[src]));
            f[dst] = fmaf(static_cast<float>(S::w[src]), Rn,
                          SP::d(P[src * g.ps + c]) * (1.0f / kScale<SP, NS>));
        };
        if (last == 1) static_for<S::nwall>([&](auto M) { bounce(M, std::true_type{}); });
        if (last == nl - 2) static_for<S::nwall>([&](auto M) { bounce(M, std::false_type{}); });
    }
}

// The stored values pulled to node c in its own frame, d_n = S (f_n - w_n
// R(c)), so that f_n = w_n R(c) + d_n / S; returns R(c).  Bounce-back slots
// hold values that c wrote for its wall-side neighbour, relative to that
// neighbour's density, and are moved onto R(c).  The step kernels work on
// these values: rho = R(c) + (sum d)/S and the stored result are formed from
// the small shifted magnitudes, where rebuilding f in FP32 first rounded
// away much of what the shift gains (2-D droplet, 256^2, 3,000 steps, FP16
// node shift: 3.5e-4 off FP64 that way, 6.1e-5 in this frame, FP32 storage
// 9.3e-5).
template <int Q, bool WALLS, typename SP, bool NS>
__device__ __forceinline__ float load_d(const typename SP::T* __restrict__ P, const Grid& g, int kind,
                                        const Coord& p, const Nbr& nb, int64_t c, const Ref<NS>& R,
                                        float (&d)[Q])
{
    using S = St<Q>;
    const float Rc = R(c);
    if (kind == kHoldsF) {
        static_for<Q>([&](auto I) {
            constexpr int n = decltype(I)::value;
            d[n] = SP::d(P[n * g.ps + c]);
        });
        return Rc;
    }
    static_for<Q>([&](auto I) {
        constexpr int n = decltype(I)::value;
        d[n] = SP::d(P[n * g.ps + at(nb, -S::ex[n], -S::ey[n], -S::ez[n])]);
    });
    if constexpr (WALLS) {
        const int last = (S::D == 3) ? p.z : p.y;
        const int nl = (S::D == 3) ? g.nz : g.ny;
        auto bounce = [&](auto M, auto from_dn) {
            constexpr int m = decltype(M)::value;
            constexpr bool DN = decltype(from_dn)::value;
            constexpr int src = DN ? S::wall_dn[m] : S::wall_up[m];
            constexpr int dst = DN ? S::wall_up[m] : S::wall_dn[m];
            const float Rn = R(at(nb, S::ex[src], S::ey[src], S::ez[src]));
            d[dst] = fmaf(kScale<SP, NS> * static_cast<float>(S::w[src]), Rn - Rc, SP::d(P[src * g.ps + c]));
        };
        if (last == 1) static_for<S::nwall>([&](auto M) { bounce(M, std::true_type{}); });
        if (last == nl - 2) static_for<S::nwall>([&](auto M) { bounce(M, std::false_type{}); });
    }
    return Rc;
}

__device__ __forceinline__ float psi_of(float rho) { return 1.0f - __expf(-rho); }

// Pass 1: psi at every node (from rho0 when given), and with the node shift
// the density of the gathered f (the reference density of this step's
// output).  P == nullptr: initialisation (rho0 only).
template <int Q, bool WALLS, typename SP, bool NS>
__global__ void __launch_bounds__(kBlock)
    k_psi(const typename SP::T* __restrict__ P, int kind, const double* __restrict__ rho0,
          float* __restrict__ psi, Grid g, Ref<NS> rin, float* __restrict__ rout)
{
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (c >= g.n) return;
    float rho = rho0 ? static_cast<float>(rho0[c]) : 0.f;
    float rf = rho;
    if (P && (!rho0 || NS)) {
        const Coord p = decompose(static_cast<uint32_t>(c), g);
        const Nbr nb = neighbours(p, g);
        float d[Q];
        const float Rc = load_d<Q, WALLS, SP, NS>(P, g, kind, p, nb, c, rin, d);
        float sd = 0.f;
        static_for<Q>([&](auto I) { sd += d[decltype(I)::value]; });
        rf = fmaf(sd, 1.0f / kScale<SP, NS>, Rc);
        if (!rho0) rho = rf;
    }
    psi[c] = psi_of(rho);
    if constexpr (NS) rout[c] = rf;
}

// Pass 2.  MODE 0: moments of f; 1: rho = rho0, u = 0; 2: initial
// equilibrium (written as f held at its own node).
template <int Q, bool WALLS, int MODE, typename SP, bool NS>
__global__ void __launch_bounds__(kBlock)
    k_collide(const typename SP::T* __restrict__ Pin, int kind, typename SP::T* __restrict__ Pout,
              const float* __restrict__ psi, const double* __restrict__ rho0, Grid g, float tau,
              float G, Ref<NS> rin, Ref<NS> rout)
{
    using S = St<Q>;
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (c >= g.n) return;
    const Coord p = decompose(static_cast<uint32_t>(c), g);
    const Nbr nb = neighbours(p, g);
    constexpr float Sc = kScale<SP, NS>;
    // pulled values in this node's frame (MODE 2: none, the equilibrium of rho0)
    float d[Q];
    float Rc = 0.f;
    float rho, u = 0.f, v = 0.f, w = 0.f;
    if constexpr (MODE != 2) Rc = load_d<Q, WALLS, SP, NS>(Pin, g, kind, p, nb, c, rin, d);
    if constexpr (MODE == 0) {
        float sd = 0.f, sx = 0.f, sy = 0.f, sz = 0.f;
        static_for<Q>([&](auto I) {
            constexpr int n = decltype(I)::value;
            sd += d[n];
            acc_e<S::ex[n]>(sx, d[n]);
            acc_e<S::ey[n]>(sy, d[n]);
            acc_e<S::ez[n]>(sz, d[n]);
        });
        rho = fmaf(sd, 1.0f / Sc, Rc);
        const float js = (1.0f / rho) * (1.0f / Sc);   // momentum of f equals that of d / S
        u = sx * js; v = sy * js; w = sz * js;
    } else {
        rho = static_cast<float>(rho0[c]);
    }
    const float pc = psi[c];
    float gx = 0.f, gy = 0.f, gz = 0.f;
    static_for<Q>([&](auto I) {
        constexpr int n = decltype(I)::value;
        if constexpr (n > 0) {
            const float t = static_cast<float>(S::w[n]) * psi[at(nb, S::ex[n], S::ey[n], S::ez[n])];
            acc_e<S::ex[n]>(gx, t);
            acc_e<S::ey[n]>(gy, t);
            acc_e<S::ez[n]>(gz, t);
        }
    });
    const float k = (-G * pc) * tau / rho;
    const float ueq = u + k * gx, veq = v + k * gy, weq = (S::D == 3) ? w + k * gz : 0.f;
    const float usq15 = 1.5f * (ueq * ueq + veq * veq + weq * weq);
    const float omega = 1.0f / tau, omt = 1.0f - omega;
    // stored_n = d_n omt + (S w_n rho omega) pm1_n + S w_n (omega (rho - Rc) + Rc - Rd),
    // pm1_n = poly_n - 1: S (f*_n - w_n Rd) for f*_n = f_n - (f_n - feq_n)/tau
    // and the reading node's reference density Rd
    const float rom = rho * omega, ct = omega * (rho - Rc);
    static_for<Q>([&](auto I) {
        constexpr int n = decltype(I)::value;
        constexpr float sw = Sc * static_cast<float>(S::w[n]);
        const float ev = edot<S::ex[n], S::ey[n], S::ez[n]>(ueq, veq, weq);
        const float pm1 = fmaf(fmaf(4.5f, ev, 3.0f), ev, -usq15);
        float out;
        if constexpr (MODE == 2) {
            // the equilibrium, read in place by this node (Rd = its own R)
            out = fmaf(sw * rho, pm1, sw * (rho - rout(c)));
        } else {
            const float Rd = rout(at(nb, S::ex[n], S::ey[n], S::ez[n]));
            out = fmaf(d[n], omt, fmaf(sw * rom, pm1, sw * (ct + (Rc - Rd))));
        }
        Pout[n * g.ps + c] = SP::of_d(out);
    });
}

template <int Q, bool WALLS, typename SP, bool NS>
__global__ void __launch_bounds__(kBlock)
    k_moments(const typename SP::T* __restrict__ P, int kind, Grid g, Ref<NS> rin,
              double* __restrict__ rho, double* __restrict__ u, double* __restrict__ v,
              double* __restrict__ w)
{
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (c >= g.n) return;
    const Coord p = decompose(static_cast<uint32_t>(c), g);
    const Nbr nb = neighbours(p, g);
    float f[Q];
    load_f<Q, WALLS, SP, NS>(P, g, kind, p, nb, c, rin, f);
    // FP64 sums of the rebuilt values, so the reported moments carry no
    // extra FP32 rounding.
    double r = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
    static_for<Q>([&](auto I) {
        constexpr int n = decltype(I)::value;
        const double fv = f[n];
        r += fv;
        acc_e<St<Q>::ex[n]>(sx, fv);
        acc_e<St<Q>::ey[n]>(sy, fv);
        acc_e<St<Q>::ez[n]>(sz, fv);
    });
    rho[c] = r;
    u[c] = sx / r;
    v[c] = sy / r;
    w[c] = (St<Q>::D == 3) ? sz / r : 0.0;
}

template <int Q, bool WALLS, typename SP, bool NS>
__global__ void __launch_bounds__(kBlock)
    k_get_f(const typename SP::T* __restrict__ P, int kind, Grid g, Ref<NS> rin,
            double* __restrict__ aos)
{
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (c >= g.n) return;
    const Coord p = decompose(static_cast<uint32_t>(c), g);
    const Nbr nb = neighbours(p, g);
    float f[Q];
    load_f<Q, WALLS, SP, NS>(P, g, kind, p, nb, c, rin, f);
#pragma unroll
    for (int n = 0; n < Q; n++) aos[c * Q + n] = f[n];
}

template <int Q, typename SP, bool NS>
__global__ void __launch_bounds__(kBlock)
    k_set_f(const double* __restrict__ aos, Grid g, float rs, typename SP::T* __restrict__ P,
            float* __restrict__ rout)
{
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (c >= g.n) return;
    float f[Q];
#pragma unroll
    for (int n = 0; n < Q; n++) f[n] = static_cast<float>(aos[c * Q + n]);
    float R = rs;
    if constexpr (NS) {
        R = 0.f;
        static_for<Q>([&](auto I) { R += f[decltype(I)::value]; });
        rout[c] = R;
    }
    static_for<Q>([&](auto I) {
        constexpr int n = decltype(I)::value;
        P[n * g.ps + c] = SP::of_d(fmaf(-static_cast<float>(St<Q>::w[n]), R, f[n]) * kScale<SP, NS>);
    });
}

// Fixed-shape block reduction (kConsBlock x kConsPerThread nodes per
// block, lbm_internal.h), the same blocks as the FP64 one.
template <int Q, bool WALLS, typename SP, bool NS>
__global__ void __launch_bounds__(kConsBlock)
    k_conservation(const typename SP::T* __restrict__ P, int kind, Grid g, Ref<NS> rin,
                   double* __restrict__ part)
{
    __shared__ double red[4][kConsBlock];
    double m = 0.0, jx = 0.0, jy = 0.0, jz = 0.0;
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kConsBlock * kConsPerThread;
    for (int i = 0; i < kConsPerThread; i++) {
        const int64_t c = base + static_cast<int64_t>(i) * kConsBlock + threadIdx.x;
        if (c >= g.n) break;
        const Coord p = decompose(static_cast<uint32_t>(c), g);
        const Nbr nb = neighbours(p, g);
        float f[Q];
        load_f<Q, WALLS, SP, NS>(P, g, kind, p, nb, c, rin, f);
        static_for<Q>([&](auto I) {
            constexpr int n = decltype(I)::value;
            const double fv = f[n];
            m += fv;
            acc_e<St<Q>::ex[n]>(jx, fv);
            acc_e<St<Q>::ey[n]>(jy, fv);
            acc_e<St<Q>::ez[n]>(jz, fv);
        });
    }
    red[0][threadIdx.x] = m; red[1][threadIdx.x] = jx;
    red[2][threadIdx.x] = jy; red[3][threadIdx.x] = jz;
    __syncthreads();
    for (int s = kConsBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int k = 0; k < 4; k++) red[k][threadIdx.x] += red[k][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < 4) part[blockIdx.x * 4 + threadIdx.x] = red[threadIdx.x][0];
}

inline unsigned blocks_for(int64_t n) { return static_cast<unsigned>((n + kBlock - 1) / kBlock); }

// Runs CALL with Q, W (walls), SP (storage) and NS (node shift) bound as
// compile-time names.
#define LOWP_DISPATCH(prec, a, CALL)                                                    \
    do {                                                                               \
        auto go_ = [&](auto qtag, auto wtag, auto sptag, auto nstag) {                 \
            constexpr int Q = decltype(qtag)::value;                                   \
            constexpr bool W = decltype(wtag)::value;                                  \
            constexpr bool NS = decltype(nstag)::value;                                \
            using SP = decltype(sptag);                                                \
            CALL;                                                                      \
        };                                                                             \
        auto pick_sp_ = [&](auto qtag, auto wtag) {                                    \
            switch (prec) {                                                            \
            case LBM_FP32: go_(qtag, wtag, StF32{}, std::false_type{}); break;         \
            case LBM_FP16S: go_(qtag, wtag, StF16{}, std::false_type{}); break;        \
            case LBM_BF16S: go_(qtag, wtag, StBF16{}, std::false_type{}); break;       \
            case LBM_FP16N: go_(qtag, wtag, StF16{}, std::true_type{}); break;         \
            default: go_(qtag, wtag, StBF16{}, std::true_type{}); break;               \
            }                                                                          \
        };                                                                             \
        auto pick_w_ = [&](auto qtag) {                                                \
            if ((a).walls) pick_sp_(qtag, std::true_type{});                           \
            else pick_sp_(qtag, std::false_type{});                                    \
        };                                                                             \
        if ((a).q == 9) pick_w_(std::integral_constant<int, 9>{});                     \
        else pick_w_(std::integral_constant<int, 19>{});                               \
    } while (0)

float shift_of(int precision, double rho_shift)
{
    return precision == LBM_FP32 ? 0.f : static_cast<float>(rho_shift);
}

}  // namespace

bool lowp_shape_ok(int precision, const StepArgs& a, const Grid& g, int shape)
{
    if (fused::is_pair_shape(shape))
        return lowp_pairs(precision, a) && fused::pair_supported(g, a.q, shape);
    return fused::supported(g, a.q, shape) || fused::partial_supported(g, a.q, shape);
}

// A pair tile where one applies, else a one-node tile that divides the
// lattice, else a partial one-node tile (8x64, 4x64, 8x32).
int lowp_default_shape(int precision, const StepArgs& a, const Grid& g)
{
    if (lowp_pairs(precision, a)) {
        const int s = fused::default_pair_shape(g, a.q);
        if (s >= 0) return s;
    }
    const int s = fused::default_shape(g, a.q);
    if (s >= 0) return s;
    for (int p : {3, 4, 5})
        if (fused::partial_supported(g, a.q, p)) return p;
    return -1;
}

int lowp_resolve_kernel(int precision, const StepArgs& a, const Grid& g, int requested)
{
    const bool fused_ok = lowp_default_shape(precision, a, g) >= 0;
    if (requested == LBM_KERNEL_FUSED) return fused_ok ? LBM_KERNEL_FUSED : -1;
    if (requested == LBM_KERNEL_TWO_PASS) return LBM_KERNEL_TWO_PASS;
    // automatic choice: a tile that divides the lattice; the partial tiles
    // of these forms measured slower than the two passes (GLUPS, fused vs
    // two-pass: FP16 200^3 26.5 vs 35.0, 300^3 25.6 vs 37.4; FP32 200^3
    // 20.0 vs 21.7; FP16 node shift 200^3 24.6 vs 28.6)
    int exact = lowp_pairs(precision, a) ? fused::default_pair_shape(g, a.q) : -1;
    if (exact < 0) exact = fused::default_shape(g, a.q);
    return fused::auto_fused19(g, exact, false) ? LBM_KERNEL_FUSED : LBM_KERNEL_TWO_PASS;
}

hipError_t lowp_init(int precision, const StepArgs& a, const Grid& g, double rho_shift,
                     const double* rho0, void* P, float* psi, float* rout, hipStream_t s)
{
    const float rs = shift_of(precision, rho_shift);
    LOWP_DISPATCH(precision, a, ({
        using T = typename SP::T;
        const Ref<NS> ro{rout, rs};
        k_psi<Q, W, SP, NS><<<blocks_for(g.n), kBlock, 0, s>>>(nullptr, kHoldsF, rho0, psi, g, ro,
                                                                 rout);
        k_collide<Q, W, 2, SP, NS><<<blocks_for(g.n), kBlock, 0, s>>>(
            nullptr, kHoldsF, static_cast<T*>(P), psi, rho0, g, (float)a.tau, (float)a.G, ro, ro);
    }));
    return hipGetLastError();
}

hipError_t lowp_step(int precision, const StepArgs& a, const Grid& g, double rho_shift,
                     const void* Pin, int kind, void* Pout, const double* rho0, float* psi,
                     const float* rin, float* rout, int kernel, const FusedTuning& t,
                     hipStream_t s)
{
    const float rs = shift_of(precision, rho_shift);
    if (lowp_node_shift(precision) && (!rin || !rout)) return hipErrorInvalidValue;
    if (kernel == LBM_KERNEL_FUSED && rho0 == nullptr) {
        if (!lowp_shape_ok(precision, a, g, t.shape) || t.chunk < 1) return hipErrorNotSupported;
        switch (precision) {
        case LBM_FP32: return lowp_fused_launch<0, false>(Pin, kind, Pout, g, a, rs, rin, rout, t, s);
        case LBM_FP16S: return lowp_fused_launch<1, false>(Pin, kind, Pout, g, a, rs, rin, rout, t, s);
        case LBM_BF16S: return lowp_fused_launch<2, false>(Pin, kind, Pout, g, a, rs, rin, rout, t, s);
        case LBM_FP16N: return lowp_fused_launch<1, true>(Pin, kind, Pout, g, a, rs, rin, rout, t, s);
        default: return lowp_fused_launch<2, true>(Pin, kind, Pout, g, a, rs, rin, rout, t, s);
        }
    }
    LOWP_DISPATCH(precision, a, ({
        using T = typename SP::T;
        const T* in = static_cast<const T*>(Pin);
        T* out = static_cast<T*>(Pout);
        const Ref<NS> ri{rin, rs}, ro{rout, rs};
        k_psi<Q, W, SP, NS><<<blocks_for(g.n), kBlock, 0, s>>>(in, kind, rho0, psi, g, ri, rout);
        if (rho0)
            k_collide<Q, W, 1, SP, NS><<<blocks_for(g.n), kBlock, 0, s>>>(
                in, kind, out, psi, rho0, g, (float)a.tau, (float)a.G, ri, ro);
        else
            k_collide<Q, W, 0, SP, NS><<<blocks_for(g.n), kBlock, 0, s>>>(
                in, kind, out, psi, rho0, g, (float)a.tau, (float)a.G, ri, ro);
    }));
    return hipGetLastError();
}

hipError_t lowp_moments(int precision, const StepArgs& a, const Grid& g, double rho_shift,
                        const void* P, int kind, const float* rin, double* rho, double* u,
                        double* v, double* w, hipStream_t s)
{
    const float rs = shift_of(precision, rho_shift);
    LOWP_DISPATCH(precision, a, ({
        k_moments<Q, W, SP, NS><<<blocks_for(g.n), kBlock, 0, s>>>(
            static_cast<const typename SP::T*>(P), kind, g, Ref<NS>{rin, rs}, rho, u, v, w);
    }));
    return hipGetLastError();
}

hipError_t lowp_get_f(int precision, const StepArgs& a, const Grid& g, double rho_shift,
                      const void* P, int kind, const float* rin, double* aos, hipStream_t s)
{
    const float rs = shift_of(precision, rho_shift);
    LOWP_DISPATCH(precision, a, ({
        k_get_f<Q, W, SP, NS><<<blocks_for(g.n), kBlock, 0, s>>>(
            static_cast<const typename SP::T*>(P), kind, g, Ref<NS>{rin, rs}, aos);
    }));
    return hipGetLastError();
}

hipError_t lowp_set_f(int precision, const StepArgs& a, const Grid& g, double rho_shift,
                      const double* aos, void* P, float* rout, hipStream_t s)
{
    const float rs = shift_of(precision, rho_shift);
    LOWP_DISPATCH(precision, a, ({
        k_set_f<Q, SP, NS><<<blocks_for(g.n), kBlock, 0, s>>>(aos, g, rs,
                                                               static_cast<typename SP::T*>(P), rout);
    }));
    return hipGetLastError();
}

hipError_t lowp_conservation(int precision, const StepArgs& a, const Grid& g, double rho_shift,
                             const void* P, int kind, const float* rin, double* partials,
                             int* nblocks, hipStream_t s)
{
    const int nb = static_cast<int>(conservation_blocks(g.n));
    *nblocks = nb;
    const float rs = shift_of(precision, rho_shift);
    LOWP_DISPATCH(precision, a, ({
        k_conservation<Q, W, SP, NS><<<nb, kConsBlock, 0, s>>>(
            static_cast<const typename SP::T*>(P), kind, g, Ref<NS>{rin, rs}, partials);
    }));
    return hipGetLastError();
}

}  // namespace lbm
