// lbm_fused_lp.h -- single-pass D3Q19 Shan-Chen step of the reduced-precision
// forms (FP32 arithmetic; FP32, FP16 or BF16 storage).
//
// Same march as the FP64 kernel (lbm_fused.h: a workgroup owns a TY x TZ
// tile of a y-z cross-section and walks a chunk of x planes; psi of planes
// x-1..x+1 with a one-node ring sits in LDS; the pulls of plane x+2 are in
// flight during the collision of plane x), with the arithmetic rearranged
// for FP32 and for small stored magnitudes.
//
// Stored values.  A stored population is d = f - w_n * R, where R is the
// reference density of the node that will read it:
//   constant shift (NS = false): R = rs everywhere (rs = 0 for FP32 storage,
//     which then holds f itself);
//   node shift (NS = true): R = the density that the reading node had after
//     the previous step, kept as an FP32 field (rin / rout, double-buffered
//     with the populations).
// A pull at node x therefore always rebuilds f_n = d_n + w_n * R(x) with
// its own R(x), and a node y writing population n (which streams to y + e_n)
// subtracts w_n * R'(y + e_n), the density of y + e_n after this step; that
// density is in the LDS window (it is what the window's psi is computed
// from).  With the node shift the stored magnitudes are the
// non-equilibrium and velocity parts of f plus w_n times the density step
// between neighbours, so 16-bit storage resolves densities far from any
// single shift (droplets, walls) and mass drift falls by orders of
// magnitude.  Cost: one FP32 read and one FP32 write per node (8 B next to
// the 76 B of FP16 populations).
//
// Collision, with rho = R(x) + sum d, omega = 1/tau, omt = 1 - omega,
// poly_n - 1 = 3 ev + 4.5 ev^2 - 1.5 |u_eq|^2:
//   stored_n = d_n*omt + (omega*w_n*rho)*(poly_n - 1)
//              + w_n*(omega*rho + omt*R(x)) - w_n*R'(x + e_n)
// which is f*_n - w_n*R'(x + e_n) with f*_n = f_n - (f_n - feq_n)/tau.
#pragma once

#include <type_traits>

#include "lbm_fused.h"
#include "lbm_storage.h"

namespace lbm {
namespace fused {

// Register holders of the 19 stored values of one node.
//
// Load<SP>: the values as loaded, one per register and unconverted, so a
// pull issued before the collision of the previous plane stays in flight
// until the values are used after it.
// Held<SP>: a plane waiting one march step for its collision.  16-bit
// formats pack two values per register (half the registers of floats, so
// three planes plus the ring prefetch fit the occupancy target without
// spilling); the packing happens after the collision, when the loads have
// long arrived.
// Wall variants use PkFloat for both: their bounce-back values carry a
// density correction that is not a stored value.
template <typename SP> struct Load {
    // 16-bit values zero-extended into 32-bit registers: the compiler would
    // otherwise pack pairs of them, which waits for the loads at once
    uint32_t v[19];
    template <int n> __device__ __forceinline__ void load(const typename SP::T* p)
    {
        v[n] = *reinterpret_cast<const unsigned short*>(p);
    }
    template <int n> __device__ __forceinline__ float get() const
    {
        return SP::d(__builtin_bit_cast(typename SP::T, static_cast<unsigned short>(v[n])));
    }
};
template <> struct Load<StF32> {
    float v[19];
    template <int n> __device__ __forceinline__ void load(const float* p) { v[n] = *p; }
    template <int n> __device__ __forceinline__ float get() const { return v[n]; }
};
template <typename SP> struct Held;
template <> struct Held<StF16> {
    typedef _Float16 h2 __attribute__((ext_vector_type(2)));
    h2 v[10];
    __device__ __forceinline__ void pack(const Load<StF16>& l)
    {
#pragma unroll
        for (int k = 0; k < 10; k++) {
            v[k][0] = __builtin_bit_cast(_Float16, static_cast<unsigned short>(l.v[2 * k]));
            if (2 * k + 1 < 19)
                v[k][1] = __builtin_bit_cast(_Float16, static_cast<unsigned short>(l.v[2 * k + 1]));
        }
    }
    template <int n> __device__ __forceinline__ float get() const
    {
        return static_cast<float>(v[n >> 1][n & 1]);
    }
};
template <> struct Held<StBF16> {
    uint32_t v[10];
    __device__ __forceinline__ void pack(const Load<StBF16>& l)
    {
#pragma unroll
        for (int k = 0; k < 10; k++) {
            const uint32_t lo = l.v[2 * k];
            const uint32_t hi = (2 * k + 1 < 19) ? l.v[2 * k + 1] : 0u;
            v[k] = lo | (hi << 16);
        }
    }
    template <int n> __device__ __forceinline__ float get() const
    {
        return (n & 1) ? __uint_as_float(v[n >> 1] & 0xffff0000u) : __uint_as_float(v[n >> 1] << 16);
    }
};
template <> struct Held<StF32> {
    float v[19];
    __device__ __forceinline__ void pack(const Load<StF32>& l)
    {
#pragma unroll
        for (int k = 0; k < 19; k++) v[k] = l.v[k];
    }
    template <int n> __device__ __forceinline__ float get() const { return v[n]; }
};
template <typename SP> struct PkFloat {
    float v[19];
    template <int n> __device__ __forceinline__ void load(const typename SP::T* p) { v[n] = SP::d(*p); }
    template <int n> __device__ __forceinline__ float get() const { return v[n]; }
    __device__ __forceinline__ void pack(const PkFloat& l) { *this = l; }
};
// tau = 1: the collision replaces f by feq, so a waiting plane needs only
// its momentum (the density is in the LDS window), three registers.
struct HeldJ {
    float jx, jy, jz;
    template <typename L> __device__ __forceinline__ void pack(const L& l)
    {
        using S = St<19>;
        jx = 0.f; jy = 0.f; jz = 0.f;
        static_for<19>([&](auto I) {
            constexpr int n = decltype(I)::value;
            const float dn = l.template get<n>();
            acc_e<S::ex[n]>(jx, dn);
            acc_e<S::ey[n]>(jy, dn);
            acc_e<S::ez[n]>(jz, dn);
        });
    }
};
template <typename SP, bool WALLS>
using LoadT = std::conditional_t<WALLS, PkFloat<SP>, Load<SP>>;
template <typename SP, bool WALLS, bool T1>
using HeldT = std::conditional_t<T1, HeldJ, std::conditional_t<WALLS, PkFloat<SP>, Held<SP>>>;

// Storage scale: FP16 node-shifted values are stored times 2^10.  They are
// small (non-equilibrium and velocity parts, w*(density step), below 1 in
// magnitude), so unscaled many would fall below FP16's smallest normal
// (2^-14) and lose relative precision; scaled, they keep 11 significant bits
// down to 2^-24 and stay far below FP16's largest value.  A power of two, so
// scaling rounds nothing; it is folded into constants.
template <typename SP, bool NS>
constexpr float kStoreScale = (NS && std::is_same_v<SP, StF16>) ? 1024.0f : 1.0f;

// Pulls the stored values of node (x, y, z) (wrapped coordinates), which
// are relative to the node's own reference density, and returns that
// density R(x) (rin[c] with the node shift, rs otherwise).  Bounce-back slots
// hold values written by the node itself for its wall-side neighbour, so
// with the node shift they are moved onto R(x): d += w*(R(neighbour) - R(x)).
template <typename SP, bool WALLS, bool NS>
__device__ __forceinline__ float pull_lp(const typename SP::T* __restrict__ P,
                                         const float* __restrict__ rin, float rs, const Grid& g,
                                         int kind, int x, int y, int z, LoadT<SP, WALLS>& d)
{
    using S = St<19>;
    const uint32_t c = (static_cast<uint32_t>(x) * static_cast<uint32_t>(g.ny) +
                        static_cast<uint32_t>(y)) * static_cast<uint32_t>(g.nz) +
                       static_cast<uint32_t>(z);
    float R = rs;
    if constexpr (NS) R = *elem(rin, c);
    // in place (kHoldsF): zero offsets; pulled: the upstream node x - e_n
    const NbrOff o(g, x, y, z, kind == kHoldsF ? 0u : ~0u);
    static_for<19>([&](auto I) {
        constexpr int n = decltype(I)::value;
        const uint32_t up = c + o.template to<-S::ex[n], -S::ey[n], -S::ez[n]>();
        d.template load<n>(elem(P + n * g.ps, up));
    });
    if (kind == kHoldsF) return R;
    if constexpr (WALLS) {
        auto bounce = [&](auto M, auto from_dn) {
            constexpr int m = decltype(M)::value;
            constexpr bool DN = decltype(from_dn)::value;
            constexpr int src = DN ? S::wall_dn[m] : S::wall_up[m];
            constexpr int dst = DN ? S::wall_up[m] : S::wall_dn[m];
            float v = SP::d(*elem(P + src * g.ps, c));
            if constexpr (NS) {
                const float Rn = *elem(rin, c + o.template to<S::ex[src], S::ey[src], S::ez[src]>());
                v = fmaf(kStoreScale<SP, NS> * static_cast<float>(S::w[src]), Rn - R, v);
            }
            d.v[dst] = v;
        };
        if (z == 1) static_for<S::nwall>([&](auto M) { bounce(M, std::true_type{}); });
        if (z == g.nz - 2) static_for<S::nwall>([&](auto M) { bounce(M, std::false_type{}); });
    }
    return R;
}

template <typename PK> __device__ __forceinline__ float sum19(const PK& d)
{
    float s = 0.f;
    static_for<19>([&](auto I) { s += d.template get<decltype(I)::value>(); });
    return s;
}

__device__ __forceinline__ float psi_lp(float rho) { return 1.0f - __expf(-rho); }

// Collision of one node from its stored values (or, tau = 1, its momentum)
// and the LDS window, split in two so that a thread can interleave the
// outputs of two nodes: lp_prep() forms the per-node scalars, lp_out<n>()
// the stored value of population n.  Every fused reduced-precision kernel
// calls these, so a node's result does not depend on the kernel or its
// launch configuration.
//
// LpNode: where the node sits in the window -- the slots of planes x-1, x,
// x+1, its index within a plane, and the window offsets of its z-1 / z+1
// neighbours.
struct LpNode {
    const int (&sl)[3];
    int li, dzm, dzp;
    // window index of node + e_n
    template <int n, int LZ> __device__ __forceinline__ int at() const
    {
        using S = St<19>;
        const int dz = S::ez[n] < 0 ? dzm : (S::ez[n] > 0 ? dzp : 0);
        return sl[1 + S::ex[n]] + li + S::ey[n] * LZ + dz;
    }
};
struct LpColl {
    float rho, ue, ve, we, usq15, t;
};

template <typename SP, bool NS, bool T1, int LZ, typename V>
__device__ __forceinline__ LpColl lp_prep(const V& d, const LpNode& nd, const float* wpsi,
                                          const float* wrho, float R, float tau, float G, float om,
                                          float omt)
{
    using S = St<19>;
    LpColl k;
    k.rho = wrho[nd.sl[1] + nd.li];
    // momentum of f equals that of d (sum_n w_n e_n = 0)
    float jx, jy, jz;
    if constexpr (T1) {
        jx = d.jx; jy = d.jy; jz = d.jz;
    } else {
        jx = 0.f; jy = 0.f; jz = 0.f;
        static_for<19>([&](auto I) {
            constexpr int n = decltype(I)::value;
            const float dn = d.template get<n>();
            acc_e<S::ex[n]>(jx, dn);
            acc_e<S::ey[n]>(jy, dn);
            acc_e<S::ez[n]>(jz, dn);
        });
    }
    // Shan-Chen force: sum_n w_n e_n psi(x + e_n)
    float gx = 0.f, gy = 0.f, gz = 0.f;
    static_for<19>([&](auto I) {
        constexpr int n = decltype(I)::value;
        if constexpr (n > 0) {
            const float p = static_cast<float>(S::w[n]) * wpsi[nd.template at<n, LZ>()];
            acc_e<S::ex[n]>(gx, p);
            acc_e<S::ey[n]>(gy, p);
            acc_e<S::ez[n]>(gz, p);
        }
    });
    const float ir = __builtin_amdgcn_rcpf(k.rho);
    const float jr = ir * (1.0f / kStoreScale<SP, NS>);   // j is summed from scaled values
    const float kf = (tau * ir) * ((-G) * wpsi[nd.sl[1] + nd.li]);
    k.ue = fmaf(kf, gx, jx * jr);
    k.ve = fmaf(kf, gy, jy * jr);
    k.we = fmaf(kf, gz, jz * jr);
    k.usq15 = 1.5f * fmaf(k.ue, k.ue, fmaf(k.ve, k.ve, k.we * k.we));
    k.t = fmaf(om, k.rho, omt * R);
    return k;
}

template <typename SP, bool NS, bool T1, int LZ, int n, typename V>
__device__ __forceinline__ typename SP::T lp_out(const LpColl& k, const V& d, const LpNode& nd,
                                                 const float* wrho, float rs, float om, float omt)
{
    using S = St<19>;
    constexpr float wn = static_cast<float>(S::w[n]);
    constexpr float sw = kStoreScale<SP, NS> * wn;
    const float ev = edot<S::ex[n], S::ey[n], S::ez[n]>(k.ue, k.ve, k.we);
    const float pm1 = fmaf(fmaf(4.5f, ev, 3.0f), ev, -k.usq15);
    float rnb;
    if constexpr (NS) rnb = (n == 0) ? k.rho : wrho[nd.template at<n, LZ>()];
    else rnb = rs;
    float v;
    if constexpr (T1) v = fmaf(sw * k.rho, pm1, fmaf(-sw, rnb, sw * k.rho));
    else v = fmaf(d.template get<n>(), omt, fmaf((om * sw) * k.rho, pm1, fmaf(-sw, rnb, sw * k.t)));
    return SP::of_d(v);
}

// Occupancy target: 4 wavefronts per SIMD (128 VGPRs); the wall variants
// hold their values as floats and run at 2.
template <bool WALLS> struct OccLp { static constexpr int waves = WALLS ? 2 : 4; };

// One tile of the march (TY x TZ nodes at y0, z0; planes [x0, x1)), with
// the LDS window wpsi / wrho of 4 * TL::LY * TL::LZ floats each.
//
// PARTIAL: the tile may run past the end of the y or z side (sides that are
// not multiples of the tile, ny >= TY and nz >= TZ).  Every thread then
// pulls its node wrapped periodically, so the window past the last valid
// row or column holds exactly the periodic neighbours that row or column
// needs, and only nodes inside the lattice collide and store.
template <typename SP, bool WALLS, bool NS, bool T1, typename TL, bool NTS, bool PARTIAL = false>
__device__ __forceinline__ void lp_tile(const typename SP::T* __restrict__ Pin, int kind,
                                        typename SP::T* __restrict__ Pout, const Grid& g, float tau,
                                        float G, float rs, const float* __restrict__ rin,
                                        float* __restrict__ rout, int y0, int z0, int x0, int x1,
                                        float* __restrict__ wpsi, float* __restrict__ wrho)
{
    using S = St<19>;
    using LD = LoadT<SP, WALLS>;
    using HD = HeldT<SP, WALLS, T1>;
    constexpr int TY = TL::TY, TZ = TL::TZ, LZ = TL::LZ, PL = TL::LY * TL::LZ;

    const int tid = threadIdx.x;
    const int ly = tid / TZ, lz = tid - (tid / TZ) * TZ;
    const int yi = PARTIAL ? wrap(y0 + ly, g.ny) : y0 + ly;
    const int zi = PARTIAL ? wrap(z0 + lz, g.nz) : z0 + lz;
    const bool valid = !PARTIAL || (y0 + ly < g.ny && z0 + lz < g.nz);
    constexpr int c0 = TL::ZF ? 0 : 1;
    const int li = (ly + 1) * LZ + (lz + c0);
    const int dzm = (TL::ZF && lz == 0) ? TZ - 1 : -1;
    const int dzp = (TL::ZF && lz == TZ - 1) ? 1 - TZ : 1;

    int ry[2] = {0, 0}, rz[2] = {0, 0}, lr[2] = {0, 0};
    bool has_ring[2] = {false, false};
#pragma unroll
    for (int k = 0; k < TL::RPT; k++) {
        const int r = tid + k * TL::NT;
        if (r < TL::NRING) {
            int rrow, rcol;
            ring_slot<TL>(r, rrow, rcol);
            has_ring[k] = true;
            ry[k] = wrap(y0 + rrow - 1, g.ny);
            rz[k] = wrap(z0 + rcol - c0, g.nz);
            lr[k] = rrow * LZ + rcol;
        }
    }

    auto slot = [&](int x) { return ((x - x0 + 1) & 3) * PL; };
    auto xw = [&](int x) { return x >= g.nx ? x - g.nx : (x < 0 ? x + g.nx : x); };
    const uint32_t plane = static_cast<uint32_t>(g.ny) * static_cast<uint32_t>(g.nz);
    const uint32_t cyz = static_cast<uint32_t>(yi) * static_cast<uint32_t>(g.nz) +
                         static_cast<uint32_t>(zi);

    // window entry of a pulled node: (psi, rho) with rho = R + sum d
    auto put_win = [&](int at, const LD& dv, float R) {
        const float rho = fmaf(sum19(dv), 1.0f / kStoreScale<SP, NS>, R);
        wpsi[at] = psi_lp(rho);
        wrho[at] = rho;
        return rho;
    };
    auto ring_win = [&](int xp, int s) {
#pragma unroll
        for (int k = 0; k < TL::RPT; k++) {
            if (has_ring[k]) {
                LD dr;
                const float R = pull_lp<SP, WALLS, NS>(Pin, rin, rs, g, kind, xp, ry[k], rz[k], dr);
                put_win(s + lr[k], dr, R);
            }
        }
    };
    // this thread's node of plane xu (unwrapped index); the density field
    // is written for the planes of this chunk
    auto own_win = [&](int xu, const LD& dv, float R) {
        const float rho = put_win(slot(xu) + li, dv, R);
        if constexpr (NS) {
            if (valid && xu >= x0 && xu < x1) *elem(rout, static_cast<uint32_t>(xu) * plane + cyz) = rho;
        }
    };

    HD fa, fb, fc;
    float Ra, Rb, Rc;
    {
        LD l;
        ring_win(xw(x0 - 1), slot(x0 - 1));
        Rc = pull_lp<SP, WALLS, NS>(Pin, rin, rs, g, kind, xw(x0 - 1), yi, zi, l);
        own_win(x0 - 1, l, Rc);
        ring_win(x0, slot(x0));
        Ra = pull_lp<SP, WALLS, NS>(Pin, rin, rs, g, kind, x0, yi, zi, l);
        own_win(x0, l, Ra);
        fa.pack(l);
        ring_win(xw(x0 + 1), slot(x0 + 1));
        Rb = pull_lp<SP, WALLS, NS>(Pin, rin, rs, g, kind, xw(x0 + 1), yi, zi, l);
        own_win(x0 + 1, l, Rb);
        fb.pack(l);
    }
    __syncthreads();

    const float om = 1.0f / tau, omt = 1.0f - om;

    auto collide = [&](int x, const HD& d, float R) {
        if constexpr (PARTIAL)
            if (!valid) return;
        const int sl[3] = {slot(x - 1), slot(x), slot(x + 1)};
        const LpNode nd{sl, li, dzm, dzp};
        const LpColl k = lp_prep<SP, NS, T1, LZ>(d, nd, wpsi, wrho, R, tau, G, om, omt);
        const uint32_t c = static_cast<uint32_t>(x) * plane + cyz;
        static_for<19>([&](auto I) {
            constexpr int n = decltype(I)::value;
            const typename SP::T v = lp_out<SP, NS, T1, LZ, n>(k, d, nd, wrho, rs, om, omt);
            typename SP::T* base = Pout + n * g.ps;
            if constexpr (NTS) SP::put_nt(elem(base, c), v);
            else *elem(base, c) = v;
        });
    };

    auto step = [&](int x, const HD& fcur, float Rcur, HD& fnn, float& Rnn) {
        const int x2 = xw(x + 2);
        LD ln;
        Rnn = pull_lp<SP, WALLS, NS>(Pin, rin, rs, g, kind, x2, yi, zi, ln);
        LD fr[TL::RPT];
        float Rr[TL::RPT];
#pragma unroll
        for (int k = 0; k < TL::RPT; k++)
            if (has_ring[k]) Rr[k] = pull_lp<SP, WALLS, NS>(Pin, rin, rs, g, kind, x2, ry[k], rz[k], fr[k]);
        collide(x, fcur, Rcur);
#pragma unroll
        for (int k = 0; k < TL::RPT; k++)
            if (has_ring[k]) put_win(slot(x + 2) + lr[k], fr[k], Rr[k]);
        own_win(x + 2, ln, Rnn);
        fnn.pack(ln);
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

// The one-node march kernel: one tile per workgroup.  PARTIAL: ny or nz is
// not a multiple of the tile (ny >= TY, nz >= TZ); the last tiles run as
// lp_tile's partial tiles.
template <typename SP, bool WALLS, bool NS, bool T1, typename TL, bool NTS, bool PARTIAL = false>
__global__ void __launch_bounds__(TL::NT) __attribute__((amdgpu_waves_per_eu(OccLp<WALLS>::waves)))
    k_fused_lp(const typename SP::T* __restrict__ Pin, int kind, typename SP::T* __restrict__ Pout,
               Grid g, float tau, float G, float rs, const float* __restrict__ rin,
               float* __restrict__ rout, int chunk, int order)
{
    static_assert(!(PARTIAL && TL::ZF), "full-row tiles cover nz exactly");
    constexpr int PL = TL::LY * TL::LZ;
    // window of psi and rho for planes x-1 .. x+2 (slot (x - x0 + 1) mod 4)
    __shared__ float wpsi[4 * PL], wrho[4 * PL];
    const int nty = PARTIAL ? (g.ny + TL::TY - 1) / TL::TY : g.ny / TL::TY;
    const int ntz = PARTIAL ? (g.nz + TL::TZ - 1) / TL::TZ : g.nz / TL::TZ;
    const int ncx = (g.nx + chunk - 1) / chunk;
    int ty, tz, cx;
    tile_of_block(blockIdx.x, nty, ntz, ncx, order, ty, tz, cx);
    const int z0 = tz * TL::TZ;
    const int x0 = cx * chunk;
    lp_tile<SP, WALLS, NS, T1, TL, NTS, PARTIAL>(Pin, kind, Pout, g, tau, G, rs, rin, rout,
                                                 ty * TL::TY, z0, x0, min(g.nx, x0 + chunk), wpsi,
                                                 wrho);
}

}  // namespace fused
}  // namespace lbm
