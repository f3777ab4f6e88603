// lbm_fused.h -- single-pass D3Q19 Shan-Chen step for gfx950.
//
// One kernel does what the reference does in five phases per step
// (LBfluid3D::macros -> ghostNodesRho -> scmp3D::calculateShanChenForces ->
// collision -> ghostNodesStreaming + streaming, scmp3D.cpp:90-119): every
// population is read once and written once per step.
//
// Geometry.  A workgroup owns a (TY x TZ) tile of a y-z cross-section and
// marches along x over a chunk of CX planes (z is the contiguous axis, so a
// tile row is a coalesced 8*TZ-byte segment).  The force at plane x needs
// psi at planes x-1, x, x+1 on a one-node ring around the tile, so the
// workgroup keeps a rolling window of three psi planes in LDS.  Each march
// step (a) pulls f for plane x+1 (tile + ring nodes) from the upstream
// neighbours, sums rho and writes psi(rho) into LDS -- the tile's f stays in
// registers -- then (b) collides plane x from the registers held since the
// previous step and writes its post-collision values.  Ring nodes belong to
// neighbouring tiles; their extra reads are served by L2 when those tiles run
// on the same XCD, which the block-to-tile mapping arranges (tiles adjacent
// in y are dispatched to the same XCD).
//
// The arithmetic is the reference's, operation for operation (see
// lbm_fp64.hip for the file:line map); with CT = double, EXACT = true and
// -ffp-contract=off the results are bit-identical to the two-pass kernels
// and to the reference.  EXACT = false (FP32 forms) multiplies by
// reciprocals instead of dividing.
#pragma once

#include <type_traits>

#include "lbm_exp.h"
#include "lbm_internal.h"
#include "lbm_stencil.h"

namespace lbm {
namespace fused {

// Tile of TY rows (y) by TZ columns (z, contiguous); one thread per tile
// node.  LDS holds psi planes of (TY+2) rows including the y ring.
//
// ZF ("full rows"): TZ equals the lattice's z extent, so the tile's rows
// wrap periodically within the tile and there is no z ring at all; every
// cache line a row pull touches belongs to the tile itself.  Otherwise the
// plane carries a one-column z ring on each side, whose pulls touch cache
// lines of the z-neighbour tiles.
template <int TY_, int TZ_, bool ZF_> struct Tile {
    static constexpr int TY = TY_;
    static constexpr int TZ = TZ_;
    static constexpr bool ZF = ZF_;
    static constexpr int NT = TY * TZ;
    static constexpr int LY = TY + 2;
    static constexpr int LZ = ZF ? TZ : TZ + 2;
    static constexpr int NRING = ZF ? 2 * TZ : 2 * LZ + 2 * TY;
    static constexpr int RPT = (NRING + NT - 1) / NT;   // ring nodes per thread
    static_assert(RPT <= 2, "at most two ring nodes per thread");
    static_assert(NT % 64 == 0 && NT <= 1024, "whole wavefronts");
};

// Tile shapes instantiated for the launcher and the tuner (index = id).
constexpr int kNumShapes = 6;
constexpr int kShapeTY[kNumShapes] = {1, 2, 1, 8, 4, 8};
constexpr int kShapeTZ[kNumShapes] = {512, 256, 256, 64, 64, 32};
constexpr bool kShapeZF[kNumShapes] = {true, true, true, false, false, false};

inline bool supported(const Grid& g, int q, int shape)
{
    if (q != 19 || shape < 0 || shape >= kNumShapes) return false;
    // element offsets within a population plane are 32-bit byte offsets
    // (elem() below): planes of at most 2^29 nodes, for 8-byte values
    if (g.ps >= (int64_t(1) << 29)) return false;
    // the march wraps x + 2 by one subtraction of nx: nx >= 2
    const int ty = kShapeTY[shape], tz = kShapeTZ[shape];
    if (g.ny % ty != 0 || g.nx < 2) return false;
    return kShapeZF[shape] ? g.nz == tz : g.nz % tz == 0;
}

// Default tile shape for a lattice: the first that fits in order of
// measured speed on MI355X (8x64, 4x64, 8x32, then the full-row tiles), or
// -1.  (Full-row tiles avoid the z ring but pull three rows per row, and
// measured slower.)
inline int default_shape(const Grid& g, int q)
{
    constexpr int pref[kNumShapes] = {3, 4, 5, 0, 1, 2};
    for (int s : pref)
        if (supported(g, q, s)) return s;
    return -1;
}

// Partial tiles: the z-tiled one-node shapes (3, 4, 5) on lattices whose ny
// or nz is not a multiple of the tile; the wrap of the last tile takes one
// subtraction, so ny >= TY and nz >= TZ.
inline bool partial_supported(const Grid& g, int q, int shape)
{
    if (q != 19 || shape < 0 || shape >= kNumShapes || kShapeZF[shape]) return false;
    if (g.ps >= (int64_t(1) << 29)) return false;
    return g.nx >= 2 && g.ny >= kShapeTY[shape] && g.nz >= kShapeTZ[shape];
}

// Pair tile shapes (tuning ids kNumShapes + i): two z-adjacent nodes per
// thread, TY x TZ nodes per tile, 16-bit storage only (lbm_fused_pair.h).
constexpr int kNumPairShapes = 2;
constexpr int kPairTY[kNumPairShapes] = {8, 4};
constexpr int kPairTZ[kNumPairShapes] = {128, 128};

inline bool is_pair_shape(int shape) { return shape >= kNumShapes && shape < kNumShapes + kNumPairShapes; }

inline bool pair_supported(const Grid& g, int q, int shape)
{
    if (q != 19 || !is_pair_shape(shape)) return false;
    if (g.ps >= (int64_t(1) << 29)) return false;
    const int i = shape - kNumShapes;
    return g.nx >= 2 && g.ny % kPairTY[i] == 0 && g.nz % kPairTZ[i] == 0;
}

// Default pair tile for a lattice, in order of measured speed, or -1.
inline int default_pair_shape(const Grid& g, int q)
{
    for (int i = 0; i < kNumPairShapes; i++)
        if (pair_supported(g, q, kNumShapes + i)) return kNumShapes + i;
    return -1;
}

// Default x planes per workgroup for a D3Q19 tile or pair shape.  Chunk
// counts ncx from ceil(nx / max_chunk) up to nx are scored by a time model,
// rounds * (planes per chunk + 3 prologue planes), where a round is one
// set of workgroups resident at once (256 CUs x waves_per_cu / waves per
// workgroup); chunks are balanced.  It gives max_chunk at 256^3 and 512^3
// (the measured best) and short chunks where the tiles alone would leave
// CUs idle: a fixed 128 launched 32 workgroups at 128^3 (FP64 3.5 GLUPS
// against 14.9 with chunk 16), and a plain "at least 256 workgroups" rule
// put 192^3 into a second round of 32.
inline int default_chunk(const Grid& g, int shape, int max_chunk, int waves_per_cu)
{
    int ty, tz, threads;
    if (is_pair_shape(shape)) {
        ty = kPairTY[shape - kNumShapes];
        tz = kPairTZ[shape - kNumShapes];
        threads = ty * tz / 2;
    } else if (shape >= 0 && shape < kNumShapes) {
        ty = kShapeTY[shape];
        tz = kShapeTZ[shape];
        threads = ty * tz;
    } else {
        return max_chunk;
    }
    const int64_t tiles = static_cast<int64_t>((g.ny + ty - 1) / ty) * ((g.nz + tz - 1) / tz);
    const int64_t wg_waves = (threads + 63) / 64;
    const int64_t slots = 256 * (waves_per_cu / wg_waves > 0 ? waves_per_cu / wg_waves : 1);
    int64_t best_ncx = (g.nx + max_chunk - 1) / max_chunk, best_cost = -1;
    for (int64_t ncx = best_ncx; ncx <= g.nx; ncx++) {
        const int64_t chunk = (g.nx + ncx - 1) / ncx;
        const int64_t rounds = (tiles * ncx + slots - 1) / slots;
        const int64_t cost = rounds * (chunk + 3);
        if (best_cost < 0 || cost < best_cost) {
            best_cost = cost;
            best_ncx = ncx;
        }
    }
    return static_cast<int>((g.nx + best_ncx - 1) / best_ncx);
}

// Automatic kernel choice for D3Q19: the fused kernel from 2^21 nodes; in
// the FP32-compute forms not on the one-node 8x32 tile.  Measured on MI355X
// with the default chunks (GLUPS, fused vs two-pass): FP64 64^3 10.9 vs
// 12.0, 96^3 11.2 vs 12.7, 128^3 14.9 vs 10.4, 160^3 (8x32) 13.6 vs 11.0,
// 192^3 15.4 vs 11.3; FP16 96^3 22.3 vs 32.9, 128^3 (pair) 41.0 vs 40.6,
// 160^3 (8x32) 30.3 vs 36.5; FP32 96^3 15.6 vs 23.8, 128^3 26.5 vs 24.0.
constexpr int64_t kAutoMinNodes19 = int64_t(1) << 21;

inline bool auto_fused19(const Grid& g, int shape, bool fp64)
{
    return shape >= 0 && g.n >= kAutoMinNodes19 && (fp64 || shape != 5);
}

// D2Q9 row-segment shapes (tuning ids kNumShapes + kNumPairShapes + i): a
// workgroup owns TW consecutive nodes of a y row (y is the contiguous axis
// in 2-D) and marches along x.  Any ny: a segment that runs past the row
// end wraps periodically and its extra lanes store nothing.
constexpr int kNumRowShapes = 3;
constexpr int kRowTW[kNumRowShapes] = {256, 128, 64};
constexpr int kFirstRowShape = kNumShapes + kNumPairShapes;

inline bool is_row_shape(int shape)
{
    return shape >= kFirstRowShape && shape < kFirstRowShape + kNumRowShapes;
}

inline bool row_supported(const Grid& g, int q, int shape)
{
    return q == 9 && is_row_shape(shape) && g.nz == 1 && g.nx >= 2 && g.ny >= 1 &&
           g.ps < (int64_t(1) << 29);
}

// Default segment: the narrowest that covers a short row, else 256 nodes.
inline int default_row_shape(const Grid& g, int q)
{
    if (q != 9) return -1;
    const int i = g.ny <= 64 ? 2 : (g.ny <= 128 ? 1 : 0);
    return row_supported(g, q, kFirstRowShape + i) ? kFirstRowShape + i : -1;
}

// D2Q9 FP64 temporal blocks (tuning ids kFirstBlockShape + i): K steps per
// launch on T x T tiles of a periodic lattice without walls, the tile and a
// halo of 2K nodes held in LDS.  For lattices that fit the memory-side cache
// the single-step kernels wait on one dependent load -> psi -> collide ->
// store chain per step; a block pays one load and one store per K steps.
constexpr int kNumBlockShapes = 3;
constexpr int kBlockT[kNumBlockShapes] = {16, 16, 8};
constexpr int kBlockK[kNumBlockShapes] = {2, 4, 4};
constexpr int kFirstBlockShape = kFirstRowShape + kNumRowShapes;

inline bool is_block_shape(int shape)
{
    return shape >= kFirstBlockShape && shape < kFirstBlockShape + kNumBlockShapes;
}

inline int block_steps(int shape) { return is_block_shape(shape) ? kBlockK[shape - kFirstBlockShape] : 1; }

inline bool block_supported(const Grid& g, int q, bool walls, int shape)
{
    if (q != 9 || walls || !is_block_shape(shape) || g.nz != 1) return false;
    const int t = kBlockT[shape - kFirstBlockShape];
    return g.nx % t == 0 && g.ny % t == 0 && g.n < (int64_t(1) << 30) && g.ps < (int64_t(1) << 30);
}

// Default temporal block of a periodic D2Q9 FP64 lattice (-1: none).
// Measured on MI355X, tau = 1, us per step with single steps vs the block
// taken: 128^2 4.26 vs 2.31 (8x8 x 4 steps), 256^2 7.06 vs 3.51 (16x16 x 4),
// 512^2 12.56 vs 10.35 (16x16 x 2), 1024^2 38.1 vs 32.6 (16x16 x 2); at
// 2048^2 the row march is as fast (126 vs 124), so larger lattices keep it.
inline int default_block(const Grid& g, int q, bool walls)
{
    auto fits = [&](int i) {
        return block_supported(g, q, walls, kFirstBlockShape + i) ? kFirstBlockShape + i : -1;
    };
    if (g.n <= (int64_t(1) << 14)) return fits(2) >= 0 ? fits(2) : fits(1);
    if (g.n <= (int64_t(1) << 16)) return fits(1) >= 0 ? fits(1) : fits(2);
    if (g.n <= (int64_t(1) << 20)) return fits(0);
    return -1;
}

// Default x rows per workgroup for a row segment: nx/256 in [1, 32] (about
// 256 row chunks), raised to 32 where that still gives 2048 workgroups.  A
// chunk re-pulls two boundary rows, so long chunks save traffic once the
// launch is wide enough.  Measured on MI355X (square lattices, 256-node
// segments) within 3% of the best chunk from 512^2 to 8192^2; fixed
// workgroup-count targets were not (8192^2 lost 13% at chunk 64).
inline int default_row_chunk(const Grid& g, int shape)
{
    if (!is_row_shape(shape)) return 64;
    const int tw = kRowTW[shape - kFirstRowShape];
    const int64_t nty = (g.ny + tw - 1) / tw;
    int chunk = g.nx / 256 < 1 ? 1 : (g.nx / 256 > 32 ? 32 : g.nx / 256);
    if (nty * ((g.nx + 31) / 32) >= 2048) chunk = 32;
    return chunk;
}

// Smallest D2Q9 lattice the automatic kernel choice runs fused, unless its
// rows are full segments (row_full): below it the march's dependent load
// rounds, ring included, cost more than the second pass (MI355X, FP64:
// 256^2 two-pass 8.8 vs fused 7.3 GLUPS; 512^2 18.2 vs 19.5).
constexpr int64_t kRowAutoMinNodes = int64_t(1) << 18;

// Rows that are exactly one segment (ny == segment width): the kernel's
// full-row form, with no ring.  Measured faster than the two passes at any
// size (128^2: 4.3 vs 6.8 us per step; 256^2: 7.0-7.8 vs 8.1).
inline bool row_full(const Grid& g, int shape)
{
    return is_row_shape(shape) && g.ny == kRowTW[shape - kFirstRowShape];
}

// Ring slot r -> (row, col) in LDS coordinates.  Rows 0 and LY-1 are the y
// ring; without ZF, columns 0 and LZ-1 are the z ring.
template <typename TL> __device__ __forceinline__ void ring_slot(int r, int& row, int& col)
{
    if (TL::ZF) {
        if (r < TL::TZ) { row = 0; col = r; }
        else { row = TL::LY - 1; col = r - TL::TZ; }
        return;
    }
    if (r < TL::LZ) { row = 0; col = r; }
    else if (r < 2 * TL::LZ) { row = TL::LY - 1; col = r - TL::LZ; }
    else if (r < 2 * TL::LZ + TL::TY) { row = 1 + (r - 2 * TL::LZ); col = 0; }
    else { row = 1 + (r - 2 * TL::LZ - TL::TY); col = TL::LZ - 1; }
}

__device__ __forceinline__ int wrap(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

// Element i of an array addressed as a 32-bit byte offset from a base that is
// uniform across the wavefront: the access then takes the global "saddr" form
// (64-bit base in SGPRs plus one 32-bit VGPR), with no 64-bit address
// arithmetic per gather.  Callers keep i * sizeof(T) below 2^32 (supported()).
template <typename T> __device__ __forceinline__ T* elem(T* base, uint32_t i)
{
    using B = std::conditional_t<std::is_const_v<T>, const char, char>;
    return reinterpret_cast<T*>(reinterpret_cast<B*>(base) + i * static_cast<uint32_t>(sizeof(T)));
}

// 32-bit offsets from node c = (x, y, z) to its +-1 neighbours along each
// axis, periodic.  With keep = 0 (a buffer read in place) every offset is 0,
// so in-place reads and pulls share one address path: a merged branch would
// turn each gather's 32-bit offset into a 64-bit one (no SGPR-base address
// form, two extra VALU operations and VGPRs per gather).
struct NbrOff {
    uint32_t xm, xp, ym, yp, zm, zp;
    __device__ __forceinline__ NbrOff(const Grid& g, int x, int y, int z, uint32_t keep = ~0u)
    {
        const uint32_t ny = static_cast<uint32_t>(g.ny), nz = static_cast<uint32_t>(g.nz);
        const uint32_t plane = ny * nz;
        xm = keep & ((x == 0) ? (static_cast<uint32_t>(g.nx) - 1u) * plane : 0u - plane);
        xp = keep & ((x == g.nx - 1) ? 0u - (static_cast<uint32_t>(g.nx) - 1u) * plane : plane);
        ym = keep & ((y == 0) ? (ny - 1u) * nz : 0u - nz);
        yp = keep & ((y == g.ny - 1) ? 0u - (ny - 1u) * nz : nz);
        zm = keep & ((z == 0) ? nz - 1u : 0u - 1u);
        zp = keep & ((z == g.nz - 1) ? 0u - (nz - 1u) : 1u);
    }
    // offset of node c + (EX, EY, EZ)
    template <int EX, int EY, int EZ> __device__ __forceinline__ uint32_t to() const
    {
        uint32_t o = 0u;
        if constexpr (EX == 1) o += xp;
        if constexpr (EX == -1) o += xm;
        if constexpr (EY == 1) o += yp;
        if constexpr (EY == -1) o += ym;
        if constexpr (EZ == 1) o += zp;
        if constexpr (EZ == -1) o += zm;
        return o;
    }
};

// Pull the stored values for node (x, y, z) (coordinates already wrapped
// into range): f itself in the FP64/FP32 formats, d = f - w*rho_shift in the
// 16-bit ones.  Offsets are 32-bit node indices added to per-population base pointers that
// are uniform across the wavefront, so each gather needs one VGPR of address.
template <typename SP, typename CT, bool WALLS>
__device__ __forceinline__ void pull(const typename SP::T* __restrict__ P, const Grid& g, int kind,
                                     int x, int y, int z, CT (&f)[19])
{
    using S = St<19>;
    const uint32_t ny = static_cast<uint32_t>(g.ny), nz = static_cast<uint32_t>(g.nz);
    const uint32_t c = (static_cast<uint32_t>(x) * ny + static_cast<uint32_t>(y)) * nz +
                       static_cast<uint32_t>(z);
    // in place (kHoldsF): zero offsets; pulled: the upstream node x - e_n
    const NbrOff o(g, x, y, z, kind == kHoldsF ? 0u : ~0u);
    static_for<19>([&](auto I) {
        constexpr int n = decltype(I)::value;
        const uint32_t up = c + o.template to<-S::ex[n], -S::ey[n], -S::ez[n]>();
        f[n] = SP::d(*elem(P + n * g.ps, up));
    });
    if (kind == kHoldsF) return;
    if constexpr (WALLS) {
        if (z == 1) {
            static_for<S::nwall>([&](auto I) {
                constexpr int m = decltype(I)::value;
                f[S::wall_up[m]] = SP::d(*elem(P + S::wall_dn[m] * g.ps, c));
            });
        }
        if (z == g.nz - 2) {
            static_for<S::nwall>([&](auto I) {
                constexpr int m = decltype(I)::value;
                f[S::wall_dn[m]] = SP::d(*elem(P + S::wall_up[m] * g.ps, c));
            });
        }
    }
}

// The exp table read from a copy in LDS (filled once per workgroup by
// fill_exp_tab): an LDS read instead of a global gather per psi.
struct ExpTabLds {
    const uint64_t* t;
    __device__ __forceinline__ uint64_t operator()(uint32_t i) const { return t[i]; }
};

__device__ __forceinline__ void fill_exp_tab(uint64_t* lds, int tid, int nthreads)
{
    for (int i = tid; i < 256; i += nthreads) lds[i] = kExpTab[i];
}

template <typename CT> struct Psi;
template <> struct Psi<double> {
    // scmp3D.cpp:187-190 with the libm-faithful exp
    static __device__ __forceinline__ double of(double rho) { return 1.0 - exp_libm(-rho); }
    static __device__ __forceinline__ double of(double rho, const uint64_t* tab)
    {
        return 1.0 - exp_libm_tab(-rho, ExpTabLds{tab});
    }
};
template <> struct Psi<float> {
    static __device__ __forceinline__ float of(float rho) { return 1.0f - __expf(-rho); }
    static __device__ __forceinline__ float of(float rho, const uint64_t*) { return of(rho); }
};

// rho of pulled values: sum f (reference order), or rs + sum d.
template <bool EXACT, typename CT>
__device__ __forceinline__ CT rho_of(const CT (&f)[19], CT rs)
{
    CT r = CT(0);
    static_for<19>([&](auto I) { r = r + f[decltype(I)::value]; });
    if constexpr (!EXACT) r = r + rs;
    return r;
}

// Collision of one node from its pulled values and the LDS psi window;
// writes the post-collision values at node c.
//
// EXACT (FP64): the reference's formulas on f, operation for operation.
// Otherwise the shifted frame d = f - w*rs (rs = rho_shift, 0 for FP32
// storage): rho = rs + sum d, j = sum d e (sum w e = 0), and the collision
// d' = d - (d - d_eq)/tau with d_eq = feq - w*rs = w*(rho*poly - rs), so no
// large f values are ever formed in FP32.
template <typename SP, typename CT, bool EXACT, bool TAU1, int LZ, bool NTS>
__device__ __forceinline__ void collide_store(typename SP::T* __restrict__ Pout, const Grid& g,
                                              uint32_t c, const CT (&f)[19],
                                              const CT* __restrict__ pm, const CT* __restrict__ p0,
                                              const CT* __restrict__ pp, int li, int dzm, int dzp,
                                              CT tau, CT G, CT rs)
{
    using S = St<19>;
    // moments, LBfluid3D.cpp:223-233
    CT rho = CT(0), sx = CT(0), sy = CT(0), sz = CT(0);
    static_for<19>([&](auto I) {
        constexpr int n = decltype(I)::value;
        rho = rho + f[n];
        acc_e<S::ex[n]>(sx, f[n]);
        acc_e<S::ey[n]>(sy, f[n]);
        acc_e<S::ez[n]>(sz, f[n]);
    });
    if constexpr (!EXACT) rho = rho + rs;
    CT u, v, w, irho = CT(0);
    if constexpr (EXACT) {
        u = sx / rho; v = sy / rho; w = sz / rho;
    } else {
        irho = CT(1) / rho;
        u = sx * irho; v = sy * irho; w = sz * irho;
    }
    // Shan-Chen force, scmp3D.cpp:152-173
    const CT pc = p0[li];
    CT gx = CT(0), gy = CT(0), gz = CT(0);
    static_for<19>([&](auto I) {
        constexpr int n = decltype(I)::value;
        if constexpr (n > 0) {
            const CT* pl = S::ex[n] < 0 ? pm : (S::ex[n] > 0 ? pp : p0);
            const int dz = S::ez[n] < 0 ? dzm : (S::ez[n] > 0 ? dzp : 0);
            const CT t = static_cast<CT>(S::w[n]) * pl[li + S::ey[n] * LZ + dz];
            acc_e<S::ex[n]>(gx, t);
            acc_e<S::ey[n]>(gy, t);
            acc_e<S::ez[n]>(gz, t);
        }
    });
    const CT mgp = (-G) * pc;
    const CT Fx = mgp * gx, Fy = mgp * gy, Fz = mgp * gz;
    CT ueq, veq, weq;
    if constexpr (EXACT) {
        ueq = u + (tau * Fx) / rho;
        veq = v + (tau * Fy) / rho;
        weq = w + (tau * Fz) / rho;
    } else {
        const CT k = tau * irho;
        ueq = u + k * Fx; veq = v + k * Fy; weq = w + k * Fz;
    }
    const CT usq15 = CT(1.5) * ((ueq * ueq + veq * veq) + weq * weq);
    const CT itau = EXACT ? CT(0) : CT(1) / tau;
    static_for<19>([&](auto I) {
        constexpr int n = decltype(I)::value;
        const CT ev = edot<S::ex[n], S::ey[n], S::ez[n]>(ueq, veq, weq);
        const CT poly = ((CT(1) + CT(3) * ev) + (CT(4.5) * ev) * ev) - usq15;
        CT out;
        if constexpr (EXACT) {
            const CT feq = poly * (rho * static_cast<CT>(S::w[n]));
            if constexpr (TAU1) out = f[n] - (f[n] - feq);   // x / 1.0 == x exactly
            else out = f[n] - (f[n] - feq) / tau;
        } else {
            const CT deq = static_cast<CT>(S::w[n]) * (rho * poly - rs);
            out = f[n] - (f[n] - deq) * itau;
        }
        if constexpr (NTS) SP::put_nt(elem(Pout + n * g.ps, c), SP::of_d(out));
        else *elem(Pout + n * g.ps, c) = SP::of_d(out);
    });
}

// Block index -> tile.  Workgroups are dispatched to the 8 XCDs round
// robin (block b runs on XCD b % 8), each XCD with its own L2; all XCDs
// share the memory-side cache.  With linear tile L = (cx*ntz + tz)*nty + ty
// and N = nty*ntz*ncx blocks, `order` picks
//   0  L = b: consecutive y-tiles on different XCDs, all XCDs on the same
//      planes at once (ring rows shared through the memory-side cache);
//   1  b = 8j + g -> L = g*(N/8) + j (needs N % 8 == 0): each XCD walks a
//      contiguous run of y-tiles (ring rows shared through its L2);
//   2  b = 8j + g -> ty = j % nty, t = (j / nty)*8 + g (needs ntz*ncx % 8
//      == 0): y-tiles of one column on one XCD, the 8 XCDs on 8 columns of
//      the same chunk.
//   3  b = 8k + g -> ty = g*(nty/8) + (k / ntz) % (nty/8), tz = k % ntz,
//      cx = k / (ntz*nty/8) (needs nty % 8 == 0): XCD g owns a band of
//      nty/8 y-tiles and walks all z-tiles of the band before the next
//      chunk, so both the y and the z neighbours of a tile run on its XCD.
// Each is a bijection of [0, N) when its condition holds; otherwise order 0.
__host__ __device__ inline int effective_order(int order, int nty, int ntz, int ncx)
{
    if (order == 1 && (nty * ntz * ncx) % 8 == 0) return 1;
    if (order == 2 && nty > 1 && (ntz * ncx) % 8 == 0) return 2;
    if (order == 3 && nty % 8 == 0) return 3;
    return 0;
}

__host__ __device__ inline void tile_of_block(int b, int nty, int ntz, int ncx, int order, int& ty,
                                              int& tz, int& cx)
{
    const int n = nty * ntz * ncx;
    int t;
    switch (effective_order(order, nty, ntz, ncx)) {
    case 1: {
        const int l = (b & 7) * (n >> 3) + (b >> 3);
        ty = l % nty;
        t = l / nty;
        break;
    }
    case 2: {
        const int j = b >> 3;
        ty = j % nty;
        t = (j / nty) * 8 + (b & 7);
        break;
    }
    case 3: {
        const int k = b >> 3, band = nty >> 3;
        tz = k % ntz;
        const int r = k / ntz;
        ty = (b & 7) * band + r % band;
        cx = r / band;
        return;
    }
    default:
        ty = b % nty;
        t = b / nty;
    }
    tz = t % ntz;
    cx = t / ntz;
}

// Occupancy target: FP64 keeps three planes of 19 doubles in registers
// (2 wavefronts per SIMD); the FP32-compute forms fit 4.
template <typename CT> struct Occ { static constexpr int waves = 2; };
template <> struct Occ<float> { static constexpr int waves = 4; };

// PARTIAL: ny or nz is not a multiple of the tile (needs ny >= TY, nz >= TZ).
// The last tile of a column runs past the lattice; its threads take the
// periodically wrapped node, so LDS slot s of every tile row or column holds
// the psi of wrapped coordinate origin + s - 1, which is exactly the
// neighbour the last real node needs.  Wrapped threads store nothing.
template <typename SP, typename CT, bool WALLS, bool EXACT, bool TAU1, typename TL, bool NTS,
          bool PARTIAL = false>
__global__ void __launch_bounds__(TL::NT) __attribute__((amdgpu_waves_per_eu(Occ<CT>::waves)))
    k_fused(const typename SP::T* __restrict__ Pin, int kind, typename SP::T* __restrict__ Pout,
            Grid g, CT tau, CT G, CT rho_shift, int chunk, int order)
{
    static_assert(!(PARTIAL && TL::ZF), "full-row tiles cover nz exactly");
    constexpr int TY = TL::TY, TZ = TL::TZ, LZ = TL::LZ, PL = TL::LY * TL::LZ;
    constexpr bool kTab = std::is_same_v<CT, double>;
    __shared__ CT psi_lds[4 * PL];
    __shared__ uint64_t exp_tab[kTab ? 256 : 1];
    if constexpr (kTab) {
        fill_exp_tab(exp_tab, threadIdx.x, TL::NT);
        __syncthreads();
    }

    const int nty = PARTIAL ? (g.ny + TY - 1) / TY : g.ny / TY;
    const int ntz = PARTIAL ? (g.nz + TZ - 1) / TZ : g.nz / TZ;
    const int ncx = (g.nx + chunk - 1) / chunk;
    int ty, tz, cx;
    tile_of_block(blockIdx.x, nty, ntz, ncx, order, ty, tz, cx);
    const int y0 = ty * TY, z0 = tz * TZ;
    const int x0 = cx * chunk;
    const int x1 = min(g.nx, x0 + chunk);

    const int tid = threadIdx.x;
    const int ly = tid / TZ, lz = tid - (tid / TZ) * TZ;  // tile node of this thread
    const int yi = PARTIAL ? wrap(y0 + ly, g.ny) : y0 + ly;
    const int zi = PARTIAL ? wrap(z0 + lz, g.nz) : z0 + lz;
    const bool valid = !PARTIAL || (y0 + ly < g.ny && z0 + lz < g.nz);
    constexpr int c0 = TL::ZF ? 0 : 1;                    // LDS column of z0
    const int li = (ly + 1) * LZ + (lz + c0);             // its LDS slot
    // LDS offsets of the z-1 / z+1 neighbours (periodic within a full row)
    const int dzm = (TL::ZF && lz == 0) ? TZ - 1 : -1;
    const int dzp = (TL::ZF && lz == TZ - 1) ? 1 - TZ : 1;

    // ring nodes of this thread: slots tid and tid + NT
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

    // LDS slot of plane x: (x - x0 + 1) mod 4
    auto slot = [&](int x) { return ((x - x0 + 1) & 3) * PL; };
    auto xw = [&](int x) { return x >= g.nx ? x - g.nx : (x < 0 ? x + g.nx : x); };
    auto psi_of = [&](const CT (&v)[19]) { return Psi<CT>::of(rho_of<EXACT>(v, rho_shift), exp_tab); };
    auto ring_psi = [&](int xp, int s) {
#pragma unroll
        for (int k = 0; k < TL::RPT; k++) {
            if (has_ring[k]) {
                CT fr[19];
                pull<SP, CT, WALLS>(Pin, g, kind, xp, ry[k], rz[k], fr);
                psi_lds[s + lr[k]] = psi_of(fr);
            }
        }
    };

    CT fa[19], fb[19], fc[19];
    // prologue: psi of planes x0-1, x0, x0+1; values of x0 (fa) and x0+1 (fb)
    ring_psi(xw(x0 - 1), slot(x0 - 1));
    pull<SP, CT, WALLS>(Pin, g, kind, xw(x0 - 1), yi, zi, fc);
    psi_lds[slot(x0 - 1) + li] = psi_of(fc);
    ring_psi(x0, slot(x0));
    pull<SP, CT, WALLS>(Pin, g, kind, x0, yi, zi, fa);
    psi_lds[slot(x0) + li] = psi_of(fa);
    ring_psi(xw(x0 + 1), slot(x0 + 1));
    pull<SP, CT, WALLS>(Pin, g, kind, xw(x0 + 1), yi, zi, fb);
    psi_lds[slot(x0 + 1) + li] = psi_of(fb);
    __syncthreads();

    auto step = [&](int x, CT (&fcur)[19], CT (&fnn)[19]) {
        const int x2 = xw(x + 2);
        // pulls of plane x+2, in flight during the collision of plane x
        pull<SP, CT, WALLS>(Pin, g, kind, x2, yi, zi, fnn);
        CT fr[TL::RPT][19];
#pragma unroll
        for (int k = 0; k < TL::RPT; k++)
            if (has_ring[k]) pull<SP, CT, WALLS>(Pin, g, kind, x2, ry[k], rz[k], fr[k]);
        const uint32_t c = (static_cast<uint32_t>(x) * static_cast<uint32_t>(g.ny) +
                            static_cast<uint32_t>(yi)) * static_cast<uint32_t>(g.nz) +
                           static_cast<uint32_t>(zi);
        if (valid)
            collide_store<SP, CT, EXACT, TAU1, LZ, NTS>(Pout, g, c, fcur, psi_lds + slot(x - 1),
                                                   psi_lds + slot(x), psi_lds + slot(x + 1), li,
                                                   dzm, dzp, tau, G, rho_shift);
#pragma unroll
        for (int k = 0; k < TL::RPT; k++)
            if (has_ring[k]) psi_lds[slot(x + 2) + lr[k]] = psi_of(fr[k]);
        psi_lds[slot(x + 2) + li] = psi_of(fnn);
        __syncthreads();
    };
    // three register sets rotate: (collide, next, pulled) = (a,b,c) -> (b,c,a) -> (c,a,b)
    int x = x0;
    for (; x + 2 < x1; x += 3) {
        step(x, fa, fc);
        step(x + 1, fb, fa);
        step(x + 2, fc, fb);
    }
    if (x < x1) step(x, fa, fc);
    if (x + 1 < x1) step(x + 1, fb, fa);
}

}  // namespace fused
}  // namespace lbm
