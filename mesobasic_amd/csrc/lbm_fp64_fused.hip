// lbm_fp64_fused.hip -- FP64 single-pass kernels: the D3Q19 instantiation of
// lbm_fused.h and the D2Q9 row march below.  Compiled with
// -ffp-contract=off like lbm_fp64.hip, so their results are bit-identical to
// the two-pass kernels and to the reference.
#pragma clang fp contract(off)

#include "lbm_fused.h"
#include "lbm_storage.h"

namespace lbm {

// D3Q19: a tile that divides the lattice, else a partial tile (8x64, then
// 8x32, then 4x64) where the lattice is at least one tile wide.
int fp64_default_shape(const Grid& g, int q)
{
    if (q == 9) return fused::default_row_shape(g, q);
    const int s = fused::default_shape(g, q);
    if (s >= 0) return s;
    for (int p : {3, 5, 4})
        if (fused::partial_supported(g, q, p)) return p;
    return -1;
}

bool fp64_shape_ok(const Grid& g, int q, int shape)
{
    if (q == 9) return fused::row_supported(g, q, shape);
    return fused::supported(g, q, shape) || fused::partial_supported(g, q, shape);
}

bool fp64_fused_supported(const StepArgs& a, const Grid& g) { return fp64_default_shape(g, a.q) >= 0; }

namespace fused {

// ---- D2Q9 row march --------------------------------------------------------
//
// The D2Q9 step of LBfluid2D::updateFluid_SC (LBfluid2D.cpp:227-312) after its
// first step, with the scmp3D force (scmp3D.cpp:141-179), in one pass.  A
// workgroup owns TW consecutive nodes of a y row (node index x*ny + y, y
// contiguous) and marches along x; LDS holds psi of rows x-1 .. x+2 for the
// segment plus one node on each side.  Lanes past the row end (ny not a
// multiple of TW, or ny < TW) take the periodically wrapped node, so their
// psi is the right-hand neighbour the last real node needs; they store
// nothing.

// f of node (x, y): read in place, or pulled from x - e_n with the y-wall
// substitution of LBfluid2D.cpp:374-376 (row 1) and :382-384 (row ny-2).
template <bool WALLS>
__device__ __forceinline__ void pull9(const double* __restrict__ P, const Grid& g, int kind, int x,
                                      int y, double (&f)[9])
{
    using S = St<9>;
    const uint32_t c = static_cast<uint32_t>(x) * static_cast<uint32_t>(g.ny) + static_cast<uint32_t>(y);
    const NbrOff o(g, x, y, 0, kind == kHoldsF ? 0u : ~0u);
    static_for<9>([&](auto I) {
        constexpr int n = decltype(I)::value;
        f[n] = *elem(P + n * g.ps, c + o.template to<-S::ex[n], -S::ey[n], 0>());
    });
    if (kind == kHoldsF) return;
    if constexpr (WALLS) {
        if (y == 1) {
            static_for<S::nwall>([&](auto I) {
                constexpr int m = decltype(I)::value;
                f[S::wall_up[m]] = *elem(P + S::wall_dn[m] * g.ps, c);
            });
        }
        if (y == g.ny - 2) {
            static_for<S::nwall>([&](auto I) {
                constexpr int m = decltype(I)::value;
                f[S::wall_dn[m]] = *elem(P + S::wall_up[m] * g.ps, c);
            });
        }
    }
}

// psi(rho) of pulled values, rho summed in direction order (scmp3D.cpp:187-190);
// tab: the exp table in LDS
__device__ __forceinline__ double psi9(const double (&f)[9], const uint64_t* tab)
{
    double r = 0.0;
    static_for<9>([&](auto I) { r = r + f[decltype(I)::value]; });
    return Psi<double>::of(r, tab);
}

// Moments, force, BGK collision of one node and the store of its
// post-collision values; the operation order of k_collide<9, *, 0> in
// lbm_fp64.hip (LBfluid2D.cpp:227-312 with usq = u*u + v*v).  pm/p0/pp are
// the psi rows x-1, x, x+1, li the node's slot in them and dym/dyp the slot
// offsets of its y-1 / y+1 neighbours.
template <bool TAU1, bool NTS>
__device__ __forceinline__ void collide9(double* __restrict__ Pout, const Grid& g, uint32_t c,
                                         const double (&f)[9], const double* pm, const double* p0,
                                         const double* pp, int li, int dym, int dyp, double tau,
                                         double G)
{
    using S = St<9>;
    double rho = 0.0, sx = 0.0, sy = 0.0;
    static_for<9>([&](auto I) {
        constexpr int n = decltype(I)::value;
        rho = rho + f[n];
        acc_e<S::ex[n]>(sx, f[n]);
        acc_e<S::ey[n]>(sy, f[n]);
    });
    const double u = sx / rho, v = sy / rho;
    const double pc = p0[li];
    double gx = 0.0, gy = 0.0;
    static_for<9>([&](auto I) {
        constexpr int n = decltype(I)::value;
        if constexpr (n > 0) {
            const double* pl = S::ex[n] < 0 ? pm : (S::ex[n] > 0 ? pp : p0);
            const double t = S::w[n] * pl[li + (S::ey[n] < 0 ? dym : (S::ey[n] > 0 ? dyp : 0))];
            acc_e<S::ex[n]>(gx, t);
            acc_e<S::ey[n]>(gy, t);
        }
    });
    const double mgp = (-G) * pc;
    const double Fx = mgp * gx, Fy = mgp * gy;
    const double ueq = u + (tau * Fx) / rho;
    const double veq = v + (tau * Fy) / rho;
    const double usq15 = 1.5 * (ueq * ueq + veq * veq);
    static_for<9>([&](auto I) {
        constexpr int n = decltype(I)::value;
        const double ev = edot<S::ex[n], S::ey[n], 0>(ueq, veq, 0.0);
        const double feq = (((1.0 + 3.0 * ev) + (4.5 * ev) * ev) - usq15) * (rho * S::w[n]);
        double out;
        if constexpr (TAU1) out = f[n] - (f[n] - feq);   // x / 1.0 == x exactly
        else out = f[n] - (f[n] - feq) / tau;
        if constexpr (NTS) StF64::put_nt(elem(Pout + n * g.ps, c), out);
        else *elem(Pout + n * g.ps, c) = out;
    });
}

// Block b -> segment b % nty of row chunk b / nty (nty = ceil(ny / TW)).
// The rows x0-1 .. x0+1 are pulled in a prologue; march step x pulls row
// x+2 (in flight during the collision of row x), collides row x, writes
// psi(x+2) and waits at one barrier.  Needs nx >= 2 (x+2 wraps by one
// subtraction).  The barrier is skipped only where `more` is false, which
// is uniform across the workgroup (x and x1 are).
//
// ZF (full rows, ny == TW): the segment is the whole periodic row, so there
// is no ring; the first and last lanes read their wrapped y neighbour's psi
// from the other end of the row's LDS slots, and the prologue is a single
// round of loads.
template <bool WALLS, bool TAU1, int TW, bool NTS, bool ZF>
__global__ void __launch_bounds__(TW)
    k_rows9(const double* __restrict__ Pin, int kind, double* __restrict__ Pout, Grid g, double tau,
            double G, int chunk)
{
    constexpr int PL = TW + 2;
    __shared__ double psi_lds[4 * PL];
    __shared__ uint64_t exp_tab[256];

    const int nty = (g.ny + TW - 1) / TW;
    const int ty = blockIdx.x % nty, cx = blockIdx.x / nty;
    const int y0 = ty * TW;
    const int x0 = cx * chunk;
    const int x1 = min(g.nx, x0 + chunk);
    const int tid = threadIdx.x;
    const bool valid = y0 + tid < g.ny;
    const int yi = (y0 + tid) % g.ny;
    const int li = tid + 1;
    // ring nodes: y0-1 (slot 0) on lane 0, y0+TW (slot TW+1) on the last lane
    const bool ring = !ZF && (tid == 0 || tid == TW - 1);
    // LDS slot offsets of the y-1 / y+1 neighbours (wrapped within a full row)
    const int dym = (ZF && tid == 0) ? TW - 1 : -1;
    const int dyp = (ZF && tid == TW - 1) ? 1 - TW : 1;
    const int ry = tid == 0 ? (y0 == 0 ? g.ny - 1 : y0 - 1) : (y0 + TW) % g.ny;
    const int lr = tid == 0 ? 0 : TW + 1;

    auto slot = [&](int x) { return ((x - x0 + 1) & 3) * PL; };
    auto xw = [&](int x) { return x >= g.nx ? x - g.nx : (x < 0 ? x + g.nx : x); };

    // prologue: the three own rows' pulls are issued before the (divergent)
    // ring block, so all loads are in flight together; a ring block between
    // them kept the compiler from hoisting the later pulls and serialised
    // six load -> exp round trips, which dominates short chunks
    double fa[9], fb[9], fc[9];
    const int xm = xw(x0 - 1), xp = xw(x0 + 1);
    fill_exp_tab(exp_tab, tid, TW);
    pull9<WALLS>(Pin, g, kind, xm, yi, fc);
    pull9<WALLS>(Pin, g, kind, x0, yi, fa);
    pull9<WALLS>(Pin, g, kind, xp, yi, fb);
    __syncthreads();   // exp table in LDS
    if (ring) {
        double r0[9], r1[9], r2[9];
        pull9<WALLS>(Pin, g, kind, xm, ry, r0);
        pull9<WALLS>(Pin, g, kind, x0, ry, r1);
        pull9<WALLS>(Pin, g, kind, xp, ry, r2);
        psi_lds[slot(x0 - 1) + lr] = psi9(r0, exp_tab);
        psi_lds[slot(x0) + lr] = psi9(r1, exp_tab);
        psi_lds[slot(x0 + 1) + lr] = psi9(r2, exp_tab);
    }
    psi_lds[slot(x0 - 1) + li] = psi9(fc, exp_tab);
    psi_lds[slot(x0) + li] = psi9(fa, exp_tab);
    psi_lds[slot(x0 + 1) + li] = psi9(fb, exp_tab);
    __syncthreads();

    // Row x+2 is needed while x+2 <= x1 (psi of row x1 enters the force of
    // row x1-1); the last step of a chunk pulls nothing, so a one-row chunk
    // is a single round of loads.
    auto step = [&](int x, double (&fcur)[9], double (&fnn)[9]) {
        const bool more = x + 2 <= x1;
        const int x2 = xw(x + 2);
        double fr[9];
        if (more) {
            pull9<WALLS>(Pin, g, kind, x2, yi, fnn);
            if (ring) pull9<WALLS>(Pin, g, kind, x2, ry, fr);
        }
        if (valid) {
            const uint32_t c = static_cast<uint32_t>(x) * static_cast<uint32_t>(g.ny) +
                               static_cast<uint32_t>(yi);
            collide9<TAU1, NTS>(Pout, g, c, fcur, psi_lds + slot(x - 1), psi_lds + slot(x),
                                psi_lds + slot(x + 1), li, dym, dyp, tau, G);
        }
        if (more) {
            if (ring) psi_lds[slot(x + 2) + lr] = psi9(fr, exp_tab);
            psi_lds[slot(x + 2) + li] = psi9(fnn, exp_tab);
            __syncthreads();
        }
    };
    int x = x0;
    for (; x + 2 < x1; x += 3) {
        step(x, fa, fc);
        step(x + 1, fb, fa);
        step(x + 2, fc, fb);
    }
    if (x < x1) step(x, fa, fc);
    if (x + 1 < x1) step(x + 1, fb, fa);
}


// ---- D2Q9 temporal blocks --------------------------------------------------
//
// K steps of the D2Q9 step above (after the first step, post-collision
// values in, no walls) per launch.  A workgroup owns a T x T tile and loads
// the post-collision values of the tile plus a halo of H = 2K nodes (wrapped
// periodically) into LDS once: a step needs f pulled from radius 1 and psi
// of the neighbours, whose densities are pulled from radius 2, so the
// region whose values are valid shrinks by 2 per step and the tile is what
// is left after K.  Each step: every node of the region pulls f from LDS,
// sums rho and writes psi (barrier), then every node one further in forms
// the force from the psi window and collides into its own LDS slots (the
// pulls all happened before the barrier, so the update is in place;
// barrier).  Per node the arithmetic is collide9's, so K steps here give
// the bits of K single steps.  One thread per node of the region, W = T+2H
// <= 32.
template <int T, int K, bool TAU1>
__global__ void __launch_bounds__(((T + 4 * K) * (T + 4 * K) + 63) / 64 * 64)
    k_block9(const double* __restrict__ Pin, double* __restrict__ Pout, Grid g, double tau, double G)
{
    using S = St<9>;
    constexpr int H = 2 * K, W = T + 2 * H, W2 = W * W;
    static_assert(W <= 32, "one thread per region node");
    __shared__ double fs[9 * W2];
    __shared__ double pw[W2];
    __shared__ uint64_t exp_tab[256];

    const int nby = g.ny / T;
    const int bx = blockIdx.x / nby, by = blockIdx.x - bx * nby;
    const int tid = threadIdx.x;
    const bool live = tid < W2;
    const int i = tid / W, j = tid - (tid / W) * W;
    // global node of region slot (i, j), wrapped (a region wider than the
    // lattice holds some nodes twice; every copy computes the same values)
    int gx = (bx * T - H + i) % g.nx;
    int gy = (by * T - H + j) % g.ny;
    gx += gx < 0 ? g.nx : 0;
    gy += gy < 0 ? g.ny : 0;
    const uint32_t c = static_cast<uint32_t>(gx) * static_cast<uint32_t>(g.ny) + static_cast<uint32_t>(gy);

    fill_exp_tab(exp_tab, tid, blockDim.x);
    if (live) {
        static_for<9>([&](auto I) {
            constexpr int n = decltype(I)::value;
            fs[n * W2 + tid] = *elem(Pin + n * g.ps, c);
        });
    }
    __syncthreads();

    const double mg = -G;
    for (int s = 0; s < K; s++) {
        const int lo = 2 * s + 1, hi = W - 2 * s - 1;
        double f[9];
        double rho = 0.0, sx = 0.0, sy = 0.0;
        const bool in_a = live && i >= lo && i < hi && j >= lo && j < hi;
        if (in_a) {
            static_for<9>([&](auto I) {
                constexpr int n = decltype(I)::value;
                f[n] = fs[n * W2 + tid - (S::ex[n] * W + S::ey[n])];
            });
            static_for<9>([&](auto I) {
                constexpr int n = decltype(I)::value;
                rho = rho + f[n];
                acc_e<S::ex[n]>(sx, f[n]);
                acc_e<S::ey[n]>(sy, f[n]);
            });
            pw[tid] = Psi<double>::of(rho, exp_tab);
        }
        __syncthreads();
        const bool in_b = live && i > lo && i < hi - 1 && j > lo && j < hi - 1;
        if (in_b) {
            const double u = sx / rho, v = sy / rho;
            const double pc = pw[tid];
            double gxs = 0.0, gys = 0.0;
            static_for<9>([&](auto I) {
                constexpr int n = decltype(I)::value;
                if constexpr (n > 0) {
                    const double t = S::w[n] * pw[tid + S::ex[n] * W + S::ey[n]];
                    acc_e<S::ex[n]>(gxs, t);
                    acc_e<S::ey[n]>(gys, t);
                }
            });
            const double mgp = mg * pc;
            const double Fx = mgp * gxs, Fy = mgp * gys;
            const double ueq = u + (tau * Fx) / rho;
            const double veq = v + (tau * Fy) / rho;
            const double usq15 = 1.5 * (ueq * ueq + veq * veq);
            static_for<9>([&](auto I) {
                constexpr int n = decltype(I)::value;
                const double ev = edot<S::ex[n], S::ey[n], 0>(ueq, veq, 0.0);
                const double feq = (((1.0 + 3.0 * ev) + (4.5 * ev) * ev) - usq15) * (rho * S::w[n]);
                double out;
                if constexpr (TAU1) out = f[n] - (f[n] - feq);
                else out = f[n] - (f[n] - feq) / tau;
                fs[n * W2 + tid] = out;
            });
        }
        __syncthreads();
    }
    if (live && i >= H && i < W - H && j >= H && j < W - H) {
        static_for<9>([&](auto I) {
            constexpr int n = decltype(I)::value;
            *elem(Pout + n * g.ps, c) = fs[n * W2 + tid];
        });
    }
}

}  // namespace fused

namespace {

template <bool W, bool T1, int TW, bool NTS>
void launch_rows(const double* Pin, int kind, double* Pout, const Grid& g, const StepArgs& a,
                 const FusedTuning& t, hipStream_t s)
{
    const int nty = (g.ny + TW - 1) / TW;
    const int ncx = (g.nx + t.chunk - 1) / t.chunk;
    const unsigned nb = static_cast<unsigned>(static_cast<int64_t>(nty) * ncx);
    if (g.ny == TW)
        fused::k_rows9<W, T1, TW, NTS, true><<<nb, TW, 0, s>>>(Pin, kind, Pout, g, a.tau, a.G, t.chunk);
    else
        fused::k_rows9<W, T1, TW, NTS, false><<<nb, TW, 0, s>>>(Pin, kind, Pout, g, a.tau, a.G, t.chunk);
}

template <bool W, bool T1, bool NTS>
void launch_row_shape(const double* Pin, int kind, double* Pout, const Grid& g, const StepArgs& a,
                      const FusedTuning& t, hipStream_t s)
{
    using namespace fused;
    switch (t.shape - kFirstRowShape) {
    case 0: launch_rows<W, T1, kRowTW[0], NTS>(Pin, kind, Pout, g, a, t, s); break;
    case 1: launch_rows<W, T1, kRowTW[1], NTS>(Pin, kind, Pout, g, a, t, s); break;
    default: launch_rows<W, T1, kRowTW[2], NTS>(Pin, kind, Pout, g, a, t, s); break;
    }
}

template <bool W, bool T1, typename TL, bool NTS>
void launch(const double* Pin, int kind, double* Pout, const Grid& g, const StepArgs& a,
            const FusedTuning& t, hipStream_t s)
{
    const int ncx = (g.nx + t.chunk - 1) / t.chunk;
    const bool partial = !TL::ZF && (g.ny % TL::TY != 0 || g.nz % TL::TZ != 0);
    const int nty = (g.ny + TL::TY - 1) / TL::TY, ntz = (g.nz + TL::TZ - 1) / TL::TZ;
    const unsigned nb = static_cast<unsigned>(static_cast<int64_t>(nty) * ntz * ncx);
    if constexpr (!TL::ZF) {
        if (partial) {
            fused::k_fused<StF64, double, W, true, T1, TL, NTS, true><<<nb, TL::NT, 0, s>>>(
                Pin, kind, Pout, g, a.tau, a.G, 0.0, t.chunk, t.order);
            return;
        }
    }
    fused::k_fused<StF64, double, W, true, T1, TL, NTS><<<nb, TL::NT, 0, s>>>(
        Pin, kind, Pout, g, a.tau, a.G, 0.0, t.chunk, t.order);
}

template <bool W, bool T1, bool NTS>
void launch_shape(const double* Pin, int kind, double* Pout, const Grid& g, const StepArgs& a,
                  const FusedTuning& t, hipStream_t s)
{
    using namespace fused;
    if (a.q == 9) {
        launch_row_shape<W, T1, NTS>(Pin, kind, Pout, g, a, t, s);
        return;
    }
    switch (t.shape) {
    case 0: launch<W, T1, Tile<kShapeTY[0], kShapeTZ[0], kShapeZF[0]>, NTS>(Pin, kind, Pout, g, a, t, s); break;
    case 1: launch<W, T1, Tile<kShapeTY[1], kShapeTZ[1], kShapeZF[1]>, NTS>(Pin, kind, Pout, g, a, t, s); break;
    case 2: launch<W, T1, Tile<kShapeTY[2], kShapeTZ[2], kShapeZF[2]>, NTS>(Pin, kind, Pout, g, a, t, s); break;
    case 3: launch<W, T1, Tile<kShapeTY[3], kShapeTZ[3], kShapeZF[3]>, NTS>(Pin, kind, Pout, g, a, t, s); break;
    case 4: launch<W, T1, Tile<kShapeTY[4], kShapeTZ[4], kShapeZF[4]>, NTS>(Pin, kind, Pout, g, a, t, s); break;
    default: launch<W, T1, Tile<kShapeTY[5], kShapeTZ[5], kShapeZF[5]>, NTS>(Pin, kind, Pout, g, a, t, s); break;
    }
}

template <bool W, bool T1>
void launch_nts(const double* Pin, int kind, double* Pout, const Grid& g, const StepArgs& a,
                const FusedTuning& t, hipStream_t s)
{
    if (t.nts) launch_shape<W, T1, true>(Pin, kind, Pout, g, a, t, s);
    else launch_shape<W, T1, false>(Pin, kind, Pout, g, a, t, s);
}

}  // namespace

hipError_t fp64_fused_step(const StepArgs& a, const Grid& g, const double* Pin, int kind,
                           double* Pout, const FusedTuning& t, hipStream_t s)
{
    if (!fp64_shape_ok(g, a.q, t.shape) || t.chunk < 1) return hipErrorNotSupported;
    const bool tau1 = (a.tau == 1.0);   // x / 1.0 == x: skip 19 divides, same bits
    if (a.walls) {
        if (tau1) launch_nts<true, true>(Pin, kind, Pout, g, a, t, s);
        else launch_nts<true, false>(Pin, kind, Pout, g, a, t, s);
    } else {
        if (tau1) launch_nts<false, true>(Pin, kind, Pout, g, a, t, s);
        else launch_nts<false, false>(Pin, kind, Pout, g, a, t, s);
    }
    return hipGetLastError();
}

template <int I, bool T1>
void launch_block(const double* Pin, double* Pout, const Grid& g, const StepArgs& a, hipStream_t s)
{
    using namespace fused;
    constexpr int T = kBlockT[I], K = kBlockK[I], W = T + 4 * K;
    const unsigned nb = static_cast<unsigned>((g.nx / T) * (g.ny / T));
    k_block9<T, K, T1><<<nb, (W * W + 63) / 64 * 64, 0, s>>>(Pin, Pout, g, a.tau, a.G);
}

template <bool T1>
void launch_block_shape(const double* Pin, double* Pout, const Grid& g, const StepArgs& a, int i, hipStream_t s)
{
    if (i == 0) launch_block<0, T1>(Pin, Pout, g, a, s);
    else if (i == 1) launch_block<1, T1>(Pin, Pout, g, a, s);
    else launch_block<2, T1>(Pin, Pout, g, a, s);
}

hipError_t fp64_block_step(const StepArgs& a, const Grid& g, const double* Pin, double* Pout, int shape,
                           hipStream_t s)
{
    using namespace fused;
    if (!block_supported(g, a.q, a.walls != 0, shape)) return hipErrorNotSupported;
    const int i = shape - kFirstBlockShape;
    if (a.tau == 1.0) launch_block_shape<true>(Pin, Pout, g, a, i, s);
    else launch_block_shape<false>(Pin, Pout, g, a, i, s);
    return hipGetLastError();
}

}  // namespace lbm

// ---- test hook (include/mesobasic_lbm_testing.h) ---------------------------

extern "C" int lbm_probe_tile_map(int nty, int ntz, int ncx, int order, int* ty, int* tz, int* cx)
{
    if (nty < 1 || ntz < 1 || ncx < 1 || !ty || !tz || !cx) return -1;
    const int nb = nty * ntz * ncx;
    for (int b = 0; b < nb; b++)
        lbm::fused::tile_of_block(b, nty, ntz, ncx, order, ty[b], tz[b], cx[b]);
    return 0;
}

// Memory skeleton of the fused D3Q19 march: the access pattern of k_fused
// (8x64 tiles of 512 threads, one node each, the same block -> tile band
// order, x chunks of 128 planes for 8-byte values and 64 for narrower ones,
// the pulls of the next plane issued before the stores of this one) with
// the arithmetic, the LDS window, the ring and the barriers taken out: every
// node loads its 19 upstream values and stores them unchanged.  Its rate is
// the ceiling of that pattern on the memory system, for the layout of this
// library (layout 0: one plane per population) or a blocked one (layout 1:
// the 19 populations of each 64-node z segment stored together).
namespace {

template <typename T, int LAYOUT>
__global__ void __launch_bounds__(512) k_probe_march(const T* __restrict__ A, T* __restrict__ B, int n,
                                                   int64_t ps, int chunk)
{
    using S = lbm::St<19>;
    constexpr int TY = 8, TZ = 64;
    const int nt = n / TY, nzt = n / TZ, ncx = (n + chunk - 1) / chunk;
    int ty, tz, cx;
    lbm::fused::tile_of_block(blockIdx.x, nt, nzt, ncx, 3, ty, tz, cx);
    const int y = ty * TY + static_cast<int>(threadIdx.x) / TZ;
    const int z = tz * TZ + static_cast<int>(threadIdx.x) % TZ;
    const int x0 = cx * chunk, x1 = min(n, x0 + chunk);
    auto at = [&](int q, int xx, int yy, int zz) -> size_t {
        if constexpr (LAYOUT == 0)
            return static_cast<size_t>(q) * ps + (static_cast<size_t>(xx) * n + yy) * n + zz;
        else
            return (((static_cast<size_t>(xx) * n + yy) * nzt + (zz >> 6)) * 19 + q) * 64 + (zz & 63);
    };
    auto wrap = [&](int v) { return v < 0 ? v + n : (v >= n ? v - n : v); };
    auto load = [&](int x, T (&v)[19]) {
        lbm::static_for<19>([&](auto I) {
            constexpr int q = decltype(I)::value;
            v[q] = A[at(q, wrap(x - S::ex[q]), wrap(y - S::ey[q]), wrap(z - S::ez[q]))];
        });
    };
    auto store = [&](int x, const T (&v)[19]) {
        lbm::static_for<19>([&](auto I) {
            constexpr int q = decltype(I)::value;
            __builtin_nontemporal_store(v[q], B + at(q, x, y, z));
        });
    };
    T a[19], b[19];
    load(x0, a);
    int x = x0;
    for (; x + 1 < x1; x += 2) {
        load(x + 1, b);
        store(x, a);
        if (x + 2 < x1) load(x + 2, a);
        store(x + 1, b);
    }
    if (x < x1) store(x, a);
}

template <typename T, int LAYOUT>
void launch_probe_march(const void* A, void* B, int n, int64_t ps)
{
    const int chunk = sizeof(T) == 8 ? 128 : 64;
    const unsigned nb = static_cast<unsigned>((n / 8) * (n / 64) * ((n + chunk - 1) / chunk));
    k_probe_march<T, LAYOUT><<<nb, 512>>>(static_cast<const T*>(A), static_cast<T*>(B), n, ps, chunk);
}

}  // namespace

extern "C" int lbm_probe_march(int n, int elem_bytes, int layout, int reps, double* ms)
{
    if (n < 64 || n % 64 != 0 || n > 1024 || reps < 1 || (layout != 0 && layout != 1) ||
        (elem_bytes != 2 && elem_bytes != 4 && elem_bytes != 8))
        return -1;
    const int64_t nodes = static_cast<int64_t>(n) * n * n;
    // planes padded as the lattice handles pad them (256 B, skewed by 512 B)
    const int64_t per = 256 / elem_bytes;
    int64_t ps = (nodes + per - 1) / per * per;
    if ((ps * elem_bytes) % 4096 == 0) ps += 512 / elem_bytes;
    const size_t bytes = static_cast<size_t>(19 * ps * elem_bytes);
    void* a = nullptr;
    void* b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess) return -2;
    if (hipMalloc(&b, bytes) != hipSuccess) { (void)hipFree(a); return -2; }
    int rc = 0;
    if (hipMemset(a, 0, bytes) != hipSuccess || hipMemset(b, 0, bytes) != hipSuccess) rc = -3;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    if (!rc) {
        // two untimed runs first; the buffers swap every run, as the lattice's do
        for (int r = -2; r < reps; r++) {
            if (r == 0) (void)hipEventRecord(e0, nullptr);
            const void* src = (r & 1) ? b : a;
            void* dst = (r & 1) ? a : b;
            if (elem_bytes == 8) {
                if (layout == 0) launch_probe_march<double, 0>(src, dst, n, ps);
                else launch_probe_march<double, 1>(src, dst, n, ps);
            } else if (elem_bytes == 4) {
                if (layout == 0) launch_probe_march<float, 0>(src, dst, n, ps);
                else launch_probe_march<float, 1>(src, dst, n, ps);
            } else {
                if (layout == 0) launch_probe_march<uint16_t, 0>(src, dst, n, ps);
                else launch_probe_march<uint16_t, 1>(src, dst, n, ps);
            }
        }
        (void)hipEventRecord(e1, nullptr);
        if (hipEventSynchronize(e1) != hipSuccess || hipGetLastError() != hipSuccess) rc = -3;
        float t = 0.f;
        (void)hipEventElapsedTime(&t, e0, e1);
        if (ms) *ms = t / reps;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    return rc;
}
