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

#include "lbm_exp.h"
#include "lbm_internal.h"
#include "lbm_stencil.h"

namespace lbm {
namespace fused {

constexpr int TY = 16;           // tile rows (y)
constexpr int TZ = 32;           // tile columns (z, contiguous)
constexpr int NT = TY * TZ;      // threads per workgroup: one per tile node
constexpr int LY = TY + 2;       // LDS plane rows incl. ring
constexpr int LZ = TZ + 2;       // LDS plane columns incl. ring
constexpr int NRING = 2 * LZ + 2 * TY;  // ring nodes per plane
constexpr int kChunkX = 32;      // planes per workgroup march

static_assert(NRING <= NT, "ring must fit in one pass of the workgroup");

inline bool supported(const Grid& g, int q)
{
    return q == 19 && g.ny % TY == 0 && g.nz % TZ == 0 && g.nx >= 1 && g.ny >= TY && g.nz >= TZ;
}

// Ring slot r -> (row, col) in LDS coordinates (row 0 and LY-1 are the
// y-ring rows, col 0 and LZ-1 the z-ring columns).
__device__ __forceinline__ void ring_slot(int r, int& row, int& col)
{
    if (r < LZ) { row = 0; col = r; }
    else if (r < 2 * LZ) { row = LY - 1; col = r - LZ; }
    else if (r < 2 * LZ + TY) { row = 1 + (r - 2 * LZ); col = 0; }
    else { row = 1 + (r - 2 * LZ - TY); col = LZ - 1; }
}

__device__ __forceinline__ int wrap(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

// Pull f for node (x, y, z) (all coordinates already wrapped into range).
// Offsets are 32-bit node indices added to per-population base pointers that
// are uniform across the wavefront, so each gather needs one VGPR of address.
template <typename SP, typename CT, bool WALLS>
__device__ __forceinline__ void pull(const typename SP::T* __restrict__ P, const Grid& g, int kind,
                                     int x, int y, int z, const CT (&shift)[19], CT (&f)[19])
{
    using S = St<19>;
    const uint32_t ny = static_cast<uint32_t>(g.ny), nz = static_cast<uint32_t>(g.nz);
    const uint32_t c = (static_cast<uint32_t>(x) * ny + static_cast<uint32_t>(y)) * nz +
                       static_cast<uint32_t>(z);
    if (kind == kHoldsF) {
        static_for<19>([&](auto I) {
            constexpr int n = decltype(I)::value;
            f[n] = SP::load(P[n * g.ps + c], shift[n]);
        });
        return;
    }
    // offsets of the x-1 / x+1 planes, y-1 / y+1 rows and z-1 / z+1 columns,
    // with periodic wrap
    const uint32_t plane = ny * nz;
    const uint32_t dxm = (x == 0) ? (static_cast<uint32_t>(g.nx) - 1u) * plane : 0u - plane;
    const uint32_t dxp = (x == g.nx - 1) ? 0u - (static_cast<uint32_t>(g.nx) - 1u) * plane : plane;
    const uint32_t dym = (y == 0) ? (ny - 1u) * nz : 0u - nz;
    const uint32_t dyp = (y == g.ny - 1) ? 0u - (ny - 1u) * nz : nz;
    const uint32_t dzm = (z == 0) ? nz - 1u : 0u - 1u;
    const uint32_t dzp = (z == g.nz - 1) ? 0u - (nz - 1u) : 1u;
    static_for<19>([&](auto I) {
        constexpr int n = decltype(I)::value;
        // upstream node x - e_n
        uint32_t o = c;
        if constexpr (S::ex[n] == 1) o += dxm;
        if constexpr (S::ex[n] == -1) o += dxp;
        if constexpr (S::ey[n] == 1) o += dym;
        if constexpr (S::ey[n] == -1) o += dyp;
        if constexpr (S::ez[n] == 1) o += dzm;
        if constexpr (S::ez[n] == -1) o += dzp;
        const typename SP::T* base = P + n * g.ps;
        f[n] = SP::load(base[o], shift[n]);
    });
    if constexpr (WALLS) {
        if (z == 1) {
            static_for<S::nwall>([&](auto I) {
                constexpr int m = decltype(I)::value;
                f[S::wall_up[m]] = SP::load(P[S::wall_dn[m] * g.ps + c], shift[S::wall_dn[m]]);
            });
        }
        if (z == g.nz - 2) {
            static_for<S::nwall>([&](auto I) {
                constexpr int m = decltype(I)::value;
                f[S::wall_dn[m]] = SP::load(P[S::wall_up[m] * g.ps + c], shift[S::wall_up[m]]);
            });
        }
    }
}

template <typename CT> struct Psi;
template <> struct Psi<double> {
    // scmp3D.cpp:187-190 with the libm-faithful exp
    static __device__ __forceinline__ double of(double rho) { return 1.0 - exp_libm(-rho); }
};
template <> struct Psi<float> {
    static __device__ __forceinline__ float of(float rho) { return 1.0f - __expf(-rho); }
};

template <typename CT>
__device__ __forceinline__ CT rho_of(const CT (&f)[19])
{
    CT r = CT(0);
    static_for<19>([&](auto I) { r = r + f[decltype(I)::value]; });
    return r;
}

// Collision of one node from its pulled f and the LDS psi window.  Writes
// the post-collision values at node c.
template <typename SP, typename CT, bool EXACT, bool TAU1>
__device__ __forceinline__ void collide_store(typename SP::T* __restrict__ Pout, const Grid& g,
                                              int64_t c, const CT (&f)[19],
                                              const CT* __restrict__ pm, const CT* __restrict__ p0,
                                              const CT* __restrict__ pp, int li, CT tau, CT G,
                                              const CT (&shift)[19])
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
            const CT t = static_cast<CT>(S::w[n]) * pl[li + S::ey[n] * LZ + S::ez[n]];
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
        const CT feq = (((CT(1) + CT(3) * ev) + (CT(4.5) * ev) * ev) - usq15) * (rho * static_cast<CT>(S::w[n]));
        CT out;
        if constexpr (TAU1) out = f[n] - (f[n] - feq);        // x / 1.0 == x exactly
        else if constexpr (EXACT) out = f[n] - (f[n] - feq) / tau;
        else out = f[n] - (f[n] - feq) * itau;
        Pout[n * g.ps + c] = SP::store(out, shift[n]);
    });
}

// Block index -> tile.  Workgroups are dispatched to the 8 XCDs round robin
// (block b runs on XCD b % 8); consecutive y-tiles of one (z, x-chunk)
// column get block indices 8 apart so that they share an L2.  With
// b = 8*j + g, (j, g) -> (ty = j % nty, t = (j / nty)*8 + g) is a bijection
// onto [0, nty) x [0, ntz*ncx) exactly when ntz*ncx is a multiple of 8;
// otherwise the plain row-major order is used.
__host__ __device__ constexpr bool xcd_remap(int nty, int ntz, int ncx)
{
    return nty > 1 && (ntz * ncx) % 8 == 0;
}

__host__ __device__ inline void tile_of_block(int b, int nty, int ntz, int ncx, int& ty, int& tz,
                                              int& cx)
{
    int t;
    if (xcd_remap(nty, ntz, ncx)) {
        const int g = b & 7, j = b >> 3;
        ty = j % nty;
        t = (j / nty) * 8 + g;
    } else {
        ty = b % nty;
        t = b / nty;
    }
    tz = t % ntz;
    cx = t / ntz;
}

template <typename SP, typename CT, bool WALLS, bool EXACT, bool TAU1>
__global__ void __launch_bounds__(NT, 1)
    k_fused(const typename SP::T* __restrict__ Pin, int kind, typename SP::T* __restrict__ Pout,
            Grid g, CT tau, CT G, CT rho_shift)
{
    __shared__ CT psi_lds[3][LY * LZ];

    const int nty = g.ny / TY, ntz = g.nz / TZ;
    const int ncx = (g.nx + kChunkX - 1) / kChunkX;
    int ty, tz, cx;
    tile_of_block(blockIdx.x, nty, ntz, ncx, ty, tz, cx);
    const int y0 = ty * TY, z0 = tz * TZ;
    const int x0 = cx * kChunkX;
    const int x1 = min(g.nx, x0 + kChunkX);

    const int tid = threadIdx.x;
    const int ly = tid / TZ, lz = tid - (tid / TZ) * TZ;  // tile node of this thread
    const int yi = y0 + ly, zi = z0 + lz;
    const int li = (ly + 1) * LZ + (lz + 1);              // its LDS slot

    // ring node of this thread (if any)
    int rrow = 0, rcol = 0, ry = 0, rz = 0;
    const bool has_ring = tid < NRING;
    if (has_ring) {
        ring_slot(tid, rrow, rcol);
        ry = wrap(y0 + rrow - 1, g.ny);
        rz = wrap(z0 + rcol - 1, g.nz);
    }

    CT shift[19];
    static_for<19>([&](auto I) {
        constexpr int n = decltype(I)::value;
        shift[n] = static_cast<CT>(St<19>::w[n]) * rho_shift;
    });

    // psi of plane xp (tile + ring) into LDS slot s; the tile's f into fdst.
    auto load_plane = [&](int xp, int s, CT (&fdst)[19]) {
        if (has_ring) {  // ring first: its registers are free again before fdst fills
            CT fr[19];
            pull<SP, CT, WALLS>(Pin, g, kind, xp, ry, rz, shift, fr);
            psi_lds[s][rrow * LZ + rcol] = Psi<CT>::of(rho_of(fr));
        }
        pull<SP, CT, WALLS>(Pin, g, kind, xp, yi, zi, shift, fdst);
        psi_lds[s][li] = Psi<CT>::of(rho_of(fdst));
    };

    CT fa[19], fb[19];
    // warm-up: planes x0-1 and x0
    load_plane(wrap(x0 - 1, g.nx), 0, fb);
    load_plane(x0, 1, fa);
    __syncthreads();

    int sm = 0, s0 = 1, sp = 2;  // LDS slots of planes x-1, x, x+1
    auto step = [&](int x, CT (&fc)[19], CT (&fn)[19]) {
        load_plane(x + 1 == g.nx ? 0 : x + 1, sp, fn);
        __syncthreads();
        const int64_t c = (static_cast<int64_t>(x) * g.ny + yi) * g.nz + zi;
        collide_store<SP, CT, EXACT, TAU1>(Pout, g, c, fc, psi_lds[sm], psi_lds[s0], psi_lds[sp], li,
                                           tau, G, shift);
        __syncthreads();
        const int t = sm; sm = s0; s0 = sp; sp = t;
    };
    int x = x0;
    for (; x + 1 < x1; x += 2) {
        step(x, fa, fb);
        step(x + 1, fb, fa);
    }
    if (x < x1) step(x, fa, fb);
}

}  // namespace fused
}  // namespace lbm
