// lbm_lowp.hip -- reduced-precision Shan-Chen step: FP32 arithmetic with
// FP32 storage of f, or FP16 / BF16 storage of the shifted populations
// d_i = f_i - w_i*rho_shift ("DDF shifting", Lehmann et al. 2022,
// "Accuracy and performance of the lattice Boltzmann method with 64-bit,
// 32-bit, and customized 16-bit number formats").  Same step structure and
// formulas as the FP64 path (lbm_fp64.hip, which cites the reference);
// only the number formats differ, so results are compared with the FP64
// oracle by tolerance, not bit for bit.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include "../../include/mesobasic_lbm.h"
#include "lbm_lowp.h"
#include "lbm_fused.h"
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

template <int Q> struct Shift {
    float s[Q];
    __device__ __forceinline__ explicit Shift(float rho_shift)
    {
#pragma unroll
        for (int n = 0; n < Q; n++) s[n] = static_cast<float>(St<Q>::w[n]) * rho_shift;
    }
};

template <int Q, bool WALLS, typename SP>
__device__ __forceinline__ void load_f(const typename SP::T* __restrict__ P, const Grid& g, int kind,
                                       const Coord& p, const Nbr& nb, int64_t c,
                                       const Shift<Q>& sh, float (&f)[Q])
{
    using S = St<Q>;
    if (kind == kHoldsF) {
#pragma unroll
        for (int n = 0; n < Q; n++) f[n] = SP::load(P[n * g.ps + c], sh.s[n]);
        return;
    }
#pragma unroll
    for (int n = 0; n < Q; n++)
        f[n] = SP::load(P[n * g.ps + at(nb, -S::ex[n], -S::ey[n], -S::ez[n])], sh.s[n]);
    if constexpr (WALLS) {
        const int last = (S::D == 3) ? p.z : p.y;
        const int nl = (S::D == 3) ? g.nz : g.ny;
        if (last == 1) {
#pragma unroll
            for (int m = 0; m < S::nwall; m++)
                f[S::wall_up[m]] = SP::load(P[S::wall_dn[m] * g.ps + c], sh.s[S::wall_dn[m]]);
        }
        if (last == nl - 2) {
#pragma unroll
            for (int m = 0; m < S::nwall; m++)
                f[S::wall_dn[m]] = SP::load(P[S::wall_up[m] * g.ps + c], sh.s[S::wall_up[m]]);
        }
    }
}

template <int Q>
__device__ __forceinline__ void moments(const float (&f)[Q], float& rho, float& sx, float& sy, float& sz)
{
    using S = St<Q>;
    rho = 0.f; sx = 0.f; sy = 0.f; sz = 0.f;
    static_for<Q>([&](auto I) {
        constexpr int n = decltype(I)::value;
        rho += f[n];
        acc_e<S::ex[n]>(sx, f[n]);
        acc_e<S::ey[n]>(sy, f[n]);
        acc_e<S::ez[n]>(sz, f[n]);
    });
}

__device__ __forceinline__ float psi_of(float rho) { return 1.0f - __expf(-rho); }

template <int Q, bool WALLS, typename SP>
__global__ void __launch_bounds__(kBlock)
    k_psi(const typename SP::T* __restrict__ P, int kind, const double* __restrict__ rho0,
          float* __restrict__ psi, Grid g, float rho_shift)
{
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (c >= g.n) return;
    if (rho0) {
        psi[c] = psi_of(static_cast<float>(rho0[c]));
        return;
    }
    const Coord p = decompose(static_cast<uint32_t>(c), g);
    const Nbr nb = neighbours(p, g);
    const Shift<Q> sh(rho_shift);
    float f[Q];
    load_f<Q, WALLS, SP>(P, g, kind, p, nb, c, sh, f);
    float rho = 0.f;
#pragma unroll
    for (int n = 0; n < Q; n++) rho += f[n];
    psi[c] = psi_of(rho);
}

// MODE 0: moments of f; 1: rho = rho0, u = 0; 2: initial equilibrium.
template <int Q, bool WALLS, int MODE, typename SP>
__global__ void __launch_bounds__(kBlock)
    k_collide(const typename SP::T* __restrict__ Pin, int kind, typename SP::T* __restrict__ Pout,
              const float* __restrict__ psi, const double* __restrict__ rho0, Grid g, float tau,
              float G, float rho_shift)
{
    using S = St<Q>;
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (c >= g.n) return;
    const Coord p = decompose(static_cast<uint32_t>(c), g);
    const Nbr nb = neighbours(p, g);
    const Shift<Q> sh(rho_shift);
    float f[Q];
    float rho, u = 0.f, v = 0.f, w = 0.f;
    if constexpr (MODE == 0) {
        load_f<Q, WALLS, SP>(Pin, g, kind, p, nb, c, sh, f);
        float sx, sy, sz;
        moments<Q>(f, rho, sx, sy, sz);
        const float ir = 1.0f / rho;
        u = sx * ir; v = sy * ir; w = sz * ir;
    } else {
        if constexpr (MODE == 1) load_f<Q, WALLS, SP>(Pin, g, kind, p, nb, c, sh, f);
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
    const float omega = 1.0f / tau;
    static_for<Q>([&](auto I) {
        constexpr int n = decltype(I)::value;
        const float ev = edot<S::ex[n], S::ey[n], S::ez[n]>(ueq, veq, weq);
        const float feq = (1.0f + 3.0f * ev + 4.5f * ev * ev - usq15) * (rho * static_cast<float>(S::w[n]));
        float out;
        if constexpr (MODE == 2) out = feq;
        else out = f[n] - (f[n] - feq) * omega;
        Pout[n * g.ps + c] = SP::store(out, sh.s[n]);
    });
}

template <int Q, bool WALLS, typename SP>
__global__ void __launch_bounds__(kBlock)
    k_moments(const typename SP::T* __restrict__ P, int kind, Grid g, float rho_shift,
              double* __restrict__ rho, double* __restrict__ u, double* __restrict__ v,
              double* __restrict__ w)
{
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (c >= g.n) return;
    const Coord p = decompose(static_cast<uint32_t>(c), g);
    const Nbr nb = neighbours(p, g);
    const Shift<Q> sh(rho_shift);
    float f[Q];
    load_f<Q, WALLS, SP>(P, g, kind, p, nb, c, sh, f);
    // FP64 sums of the stored values, so the reported moments carry no
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

template <int Q, bool WALLS, typename SP>
__global__ void __launch_bounds__(kBlock)
    k_get_f(const typename SP::T* __restrict__ P, int kind, Grid g, float rho_shift,
            double* __restrict__ aos)
{
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (c >= g.n) return;
    const Coord p = decompose(static_cast<uint32_t>(c), g);
    const Nbr nb = neighbours(p, g);
    const Shift<Q> sh(rho_shift);
    float f[Q];
    load_f<Q, WALLS, SP>(P, g, kind, p, nb, c, sh, f);
#pragma unroll
    for (int n = 0; n < Q; n++) aos[c * Q + n] = f[n];
}

template <int Q, typename SP>
__global__ void __launch_bounds__(kBlock)
    k_set_f(const double* __restrict__ aos, Grid g, float rho_shift, typename SP::T* __restrict__ P)
{
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (c >= g.n) return;
    const Shift<Q> sh(rho_shift);
#pragma unroll
    for (int n = 0; n < Q; n++) P[n * g.ps + c] = SP::store(static_cast<float>(aos[c * Q + n]), sh.s[n]);
}

constexpr int kConsBlock = 256;
constexpr int kConsPerThread = 16;

template <int Q, bool WALLS, typename SP>
__global__ void __launch_bounds__(kConsBlock)
    k_conservation(const typename SP::T* __restrict__ P, int kind, Grid g, float rho_shift,
                   double* __restrict__ part)
{
    __shared__ double red[4][kConsBlock];
    const Shift<Q> sh(rho_shift);
    double m = 0.0, jx = 0.0, jy = 0.0, jz = 0.0;
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kConsBlock * kConsPerThread;
    for (int i = 0; i < kConsPerThread; i++) {
        const int64_t c = base + static_cast<int64_t>(i) * kConsBlock + threadIdx.x;
        if (c >= g.n) break;
        const Coord p = decompose(static_cast<uint32_t>(c), g);
        const Nbr nb = neighbours(p, g);
        float f[Q];
        load_f<Q, WALLS, SP>(P, g, kind, p, nb, c, sh, f);
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

#define LOWP_DISPATCH(prec, a, CALL)                                                    \
    do {                                                                               \
        auto go_ = [&](auto qtag, auto wtag, auto sptag) {                             \
            constexpr int Q = decltype(qtag)::value;                                   \
            constexpr bool W = decltype(wtag)::value;                                  \
            using SP = decltype(sptag);                                                \
            CALL;                                                                      \
        };                                                                             \
        auto pick_sp_ = [&](auto qtag, auto wtag) {                                    \
            if ((prec) == LBM_FP32) go_(qtag, wtag, StF32{});                          \
            else if ((prec) == LBM_FP16S) go_(qtag, wtag, StF16{});                    \
            else go_(qtag, wtag, StBF16{});                                            \
        };                                                                             \
        auto pick_w_ = [&](auto qtag) {                                                \
            if ((a).walls) pick_sp_(qtag, std::true_type{});                           \
            else pick_sp_(qtag, std::false_type{});                                    \
        };                                                                             \
        if ((a).q == 9) pick_w_(std::integral_constant<int, 9>{});                     \
        else pick_w_(std::integral_constant<int, 19>{});                               \
    } while (0)

}  // namespace

int lowp_resolve_kernel(int, const StepArgs& a, const Grid& g, int requested)
{
    const bool fused_ok = fused::default_shape(g, a.q) >= 0;
    if (requested == LBM_KERNEL_FUSED) return fused_ok ? LBM_KERNEL_FUSED : -1;
    if (requested == LBM_KERNEL_TWO_PASS) return LBM_KERNEL_TWO_PASS;
    return fused_ok ? LBM_KERNEL_FUSED : LBM_KERNEL_TWO_PASS;
}

namespace {

template <typename SP, bool W, typename TL, bool NTS>
void launch_fused(const void* Pin, int kind, void* Pout, const Grid& g, const StepArgs& a, float rs,
                  const FusedTuning& t, hipStream_t s)
{
    using T = typename SP::T;
    const int ncx = (g.nx + t.chunk - 1) / t.chunk;
    const unsigned nb = static_cast<unsigned>((g.ny / TL::TY) * (g.nz / TL::TZ) * ncx);
    fused::k_fused<SP, float, W, false, false, TL, NTS><<<nb, TL::NT, 0, s>>>(
        static_cast<const T*>(Pin), kind, static_cast<T*>(Pout), g, (float)a.tau, (float)a.G, rs,
        t.chunk, t.order);
}

template <typename SP, bool W, bool NTS>
void launch_fused_shape(const void* Pin, int kind, void* Pout, const Grid& g, const StepArgs& a,
                        float rs, const FusedTuning& t, hipStream_t s)
{
    using namespace fused;
    switch (t.shape) {
    case 0: launch_fused<SP, W, Tile<kShapeTY[0], kShapeTZ[0], kShapeZF[0]>, NTS>(Pin, kind, Pout, g, a, rs, t, s); break;
    case 1: launch_fused<SP, W, Tile<kShapeTY[1], kShapeTZ[1], kShapeZF[1]>, NTS>(Pin, kind, Pout, g, a, rs, t, s); break;
    case 2: launch_fused<SP, W, Tile<kShapeTY[2], kShapeTZ[2], kShapeZF[2]>, NTS>(Pin, kind, Pout, g, a, rs, t, s); break;
    case 3: launch_fused<SP, W, Tile<kShapeTY[3], kShapeTZ[3], kShapeZF[3]>, NTS>(Pin, kind, Pout, g, a, rs, t, s); break;
    case 4: launch_fused<SP, W, Tile<kShapeTY[4], kShapeTZ[4], kShapeZF[4]>, NTS>(Pin, kind, Pout, g, a, rs, t, s); break;
    default: launch_fused<SP, W, Tile<kShapeTY[5], kShapeTZ[5], kShapeZF[5]>, NTS>(Pin, kind, Pout, g, a, rs, t, s); break;
    }
}

}  // namespace

hipError_t lowp_init(int precision, const StepArgs& a, const Grid& g, double rho_shift,
                     const double* rho0, void* P, float* psi, hipStream_t s)
{
    const float sh = static_cast<float>(rho_shift);
    LOWP_DISPATCH(precision, a, ({
        using T = typename SP::T;
        k_psi<Q, W, SP><<<blocks_for(g.n), kBlock, 0, s>>>(nullptr, kHoldsF, rho0, psi, g, sh);
        k_collide<Q, W, 2, SP><<<blocks_for(g.n), kBlock, 0, s>>>(
            nullptr, kHoldsF, static_cast<T*>(P), psi, rho0, g, (float)a.tau, (float)a.G, sh);
    }));
    return hipGetLastError();
}

hipError_t lowp_step(int precision, const StepArgs& a, const Grid& g, double rho_shift,
                     const void* Pin, int kind, void* Pout, const double* rho0, float* psi,
                     int kernel, const FusedTuning& t, hipStream_t s)
{
    const float sh = static_cast<float>(rho_shift);
    if (kernel == LBM_KERNEL_FUSED && rho0 == nullptr) {
        if (!fused::supported(g, a.q, t.shape) || t.chunk < 1) return hipErrorNotSupported;
        // FP32 stores f itself: the shifted frame then has rs = 0
        const float rs = (precision == LBM_FP32) ? 0.f : sh;
        auto go = [&](auto sptag) {
            using SP = decltype(sptag);
            if (a.walls) {
                if (t.nts) launch_fused_shape<SP, true, true>(Pin, kind, Pout, g, a, rs, t, s);
                else launch_fused_shape<SP, true, false>(Pin, kind, Pout, g, a, rs, t, s);
            } else {
                if (t.nts) launch_fused_shape<SP, false, true>(Pin, kind, Pout, g, a, rs, t, s);
                else launch_fused_shape<SP, false, false>(Pin, kind, Pout, g, a, rs, t, s);
            }
        };
        if (precision == LBM_FP32) go(StF32{});
        else if (precision == LBM_FP16S) go(StF16{});
        else go(StBF16{});
        return hipGetLastError();
    }
    LOWP_DISPATCH(precision, a, ({
        using T = typename SP::T;
        const T* in = static_cast<const T*>(Pin);
        T* out = static_cast<T*>(Pout);
        k_psi<Q, W, SP><<<blocks_for(g.n), kBlock, 0, s>>>(in, kind, rho0, psi, g, sh);
        if (rho0)
            k_collide<Q, W, 1, SP><<<blocks_for(g.n), kBlock, 0, s>>>(in, kind, out, psi, rho0, g,
                                                                        (float)a.tau, (float)a.G, sh);
        else
            k_collide<Q, W, 0, SP><<<blocks_for(g.n), kBlock, 0, s>>>(in, kind, out, psi, rho0, g,
                                                                        (float)a.tau, (float)a.G, sh);
    }));
    return hipGetLastError();
}

hipError_t lowp_moments(int precision, const StepArgs& a, const Grid& g, double rho_shift,
                        const void* P, int kind, double* rho, double* u, double* v, double* w,
                        hipStream_t s)
{
    const float sh = static_cast<float>(rho_shift);
    LOWP_DISPATCH(precision, a, ({
        k_moments<Q, W, SP><<<blocks_for(g.n), kBlock, 0, s>>>(
            static_cast<const typename SP::T*>(P), kind, g, sh, rho, u, v, w);
    }));
    return hipGetLastError();
}

hipError_t lowp_get_f(int precision, const StepArgs& a, const Grid& g, double rho_shift,
                      const void* P, int kind, double* aos, hipStream_t s)
{
    const float sh = static_cast<float>(rho_shift);
    LOWP_DISPATCH(precision, a, ({
        k_get_f<Q, W, SP><<<blocks_for(g.n), kBlock, 0, s>>>(static_cast<const typename SP::T*>(P),
                                                              kind, g, sh, aos);
    }));
    return hipGetLastError();
}

hipError_t lowp_set_f(int precision, const StepArgs& a, const Grid& g, double rho_shift,
                      const double* aos, void* P, hipStream_t s)
{
    const float sh = static_cast<float>(rho_shift);
    LOWP_DISPATCH(precision, a, ({
        k_set_f<Q, SP><<<blocks_for(g.n), kBlock, 0, s>>>(aos, g, sh, static_cast<typename SP::T*>(P));
    }));
    return hipGetLastError();
}

hipError_t lowp_conservation(int precision, const StepArgs& a, const Grid& g, double rho_shift,
                             const void* P, int kind, double* partials, int* nblocks,
                             hipStream_t s)
{
    const int64_t per = static_cast<int64_t>(kConsBlock) * kConsPerThread;
    const int nb = static_cast<int>((g.n + per - 1) / per);
    *nblocks = nb;
    const float sh = static_cast<float>(rho_shift);
    LOWP_DISPATCH(precision, a, ({
        k_conservation<Q, W, SP><<<nb, kConsBlock, 0, s>>>(static_cast<const typename SP::T*>(P),
                                                             kind, g, sh, partials);
    }));
    return hipGetLastError();
}

}  // namespace lbm
