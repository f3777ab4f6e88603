// lbm_fp64.hip -- FP64 Shan-Chen step for gfx950, bit-faithful to the
// reference CPU code (flowzario/mesoBasic, src/lattice_boltzmann).
//
// Built with -ffp-contract=off: the reference is compiled for baseline
// x86-64 (no FMA), so every product and sum rounds on its own.  The
// operation order of each formula is the reference's left-to-right order:
//   moments   LBfluid3D.cpp:223-233   rho = sum f; u = (sum f*ex)/rho
//   psi       scmp3D.cpp:189          1.0 - exp(-rho)   (exp: lbm_exp.h)
//   force     scmp3D.cpp:158-173      F = ((-G)*psi(x)) * sum (w*e)*psi(x+e)
//   feq       LBfluid3D.cpp:263-271   ueq = u + (tau*F)/rho,
//                                     feq = (((1+3ev)+(4.5ev)ev)-1.5usq)*(rho*w)
//   collide   LBfluid3D.cpp:272       f - (f - feq)/tau  (IEEE divide)
//   stream    LBfluid3D.cpp:288-302   f(x) = f*(x - e)   (pull)
//   walls     LBfluid3D.cpp:316-347, LBfluid2D.cpp:363-388
//
// Storage: populations are structure-of-arrays, and the buffer written by a
// step holds post-collision values f*(x).  The next step pulls
// f(x) = f*(x - e) with periodic wrap by index arithmetic, which replaces the
// reference's ghost shell and its MPI halo exchange (LBfluid3D.cpp:430-543).
// At a wall row the reference overwrites the streamed value with the same
// node's post-collision value of the opposite direction; the pull does the
// same substitution when it reads.
#pragma clang fp contract(off)

#include "lbm_exp.h"
#include "lbm_internal.h"
#include "lbm_stencil.h"

#include <algorithm>
#include <type_traits>

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

// Periodic neighbour coordinates along each axis: [0] = -1, [1] = 0, [2] = +1.
struct Nbr {
    int64_t xy[3][3];  // (x + dx)*ny + (y + dy), premultiplied by nz
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

// Node at offset (dx, dy, dz) in {-1,0,1}^3.
__device__ __forceinline__ int64_t at(const Nbr& nb, int dx, int dy, int dz)
{
    return nb.xy[dx + 1][dy + 1] + nb.z[dz + 1];
}

// f at node c: read in place, or pulled from the upstream neighbours with
// the wall substitution.
template <int Q, bool WALLS>
__device__ __forceinline__ void load_f(const double* __restrict__ P, const Grid& g, int kind,
                                       const Coord& p, const Nbr& nb, int64_t c, double (&f)[Q])
{
    using S = St<Q>;
    if (kind == kHoldsF) {
#pragma unroll
        for (int n = 0; n < Q; n++) f[n] = P[n * g.ps + c];
        return;
    }
#pragma unroll
    for (int n = 0; n < Q; n++) f[n] = P[n * g.ps + at(nb, -S::ex[n], -S::ey[n], -S::ez[n])];
    if constexpr (WALLS) {
        const int last = (S::D == 3) ? p.z : p.y;
        const int nl = (S::D == 3) ? g.nz : g.ny;
        if (last == 1) {
#pragma unroll
            for (int m = 0; m < S::nwall; m++) f[S::wall_up[m]] = P[S::wall_dn[m] * g.ps + c];
        }
        if (last == nl - 2) {
#pragma unroll
            for (int m = 0; m < S::nwall; m++) f[S::wall_dn[m]] = P[S::wall_up[m] * g.ps + c];
        }
    }
}

template <int Q>
__device__ __forceinline__ void moments(const double (&f)[Q], double& rho, double& sx, double& sy,
                                        double& sz)
{
    using S = St<Q>;
    rho = 0.0; sx = 0.0; sy = 0.0; sz = 0.0;
#pragma unroll
    for (int n = 0; n < Q; n++) {
        rho = rho + f[n];
        if (S::ex[n] == 1) sx = sx + f[n];
        if (S::ex[n] == -1) sx = sx - f[n];
        if (S::ey[n] == 1) sy = sy + f[n];
        if (S::ey[n] == -1) sy = sy - f[n];
        if (S::ez[n] == 1) sz = sz + f[n];
        if (S::ez[n] == -1) sz = sz - f[n];
    }
}

__device__ __forceinline__ double psi_of(double rho) { return 1.0 - exp_libm(-rho); }

__global__ void __launch_bounds__(kBlock) k_psi_rho(const double* __restrict__ rho,
                                                    double* __restrict__ psi, int64_t n)
{
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (c < n) psi[c] = psi_of(rho[c]);
}

template <int Q, bool WALLS>
__global__ void __launch_bounds__(kBlock) k_psi_pops(const double* __restrict__ P, int kind,
                                                     double* __restrict__ psi, Grid g)
{
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (c >= g.n) return;
    const Coord p = decompose(static_cast<uint32_t>(c), g);
    const Nbr nb = neighbours(p, g);
    double f[Q];
    load_f<Q, WALLS>(P, g, kind, p, nb, c, f);
    double rho = 0.0;
#pragma unroll
    for (int n = 0; n < Q; n++) rho = rho + f[n];
    psi[c] = psi_of(rho);
}

// Force + collision (+ the pull stream that produced f).  MODE 0: moments of
// f; MODE 1: rho = rho0, u = 0 (LBfluid2D's first step uses the stored
// initial arrays, LBfluid2D.cpp:238-240 after setFtoFeq); MODE 2: initial
// equilibrium (LBfluid3D::setFtoFeq, LBfluid3D.cpp:180-200).  TAU1: tau is
// exactly 1.0, so (f - feq)/tau == f - feq bit for bit and the divides go.
template <int Q, bool WALLS, int MODE, bool TAU1 = false>
__global__ void __launch_bounds__(kBlock)
    k_collide(const double* __restrict__ Pin, int kind, double* __restrict__ Pout,
              const double* __restrict__ psi, const double* __restrict__ rho0, Grid g,
              double tau, double G)
{
    using S = St<Q>;
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (c >= g.n) return;
    const Coord p = decompose(static_cast<uint32_t>(c), g);
    const Nbr nb = neighbours(p, g);

    double f[Q];
    double rho, u, v, w;
    if constexpr (MODE == 0) {
        load_f<Q, WALLS>(Pin, g, kind, p, nb, c, f);
        double sx, sy, sz;
        moments<Q>(f, rho, sx, sy, sz);
        u = sx / rho;
        v = sy / rho;
        w = sz / rho;
    } else {
        if constexpr (MODE == 1) load_f<Q, WALLS>(Pin, g, kind, p, nb, c, f);
        rho = rho0[c];
        u = 0.0; v = 0.0; w = 0.0;
    }

    // Shan-Chen force, scmp3D.cpp:141-179
    const double pc = psi[c];
    double gx = 0.0, gy = 0.0, gz = 0.0;
#pragma unroll
    for (int n = 1; n < Q; n++) {
        const double t = S::w[n] * psi[at(nb, S::ex[n], S::ey[n], S::ez[n])];
        if (S::ex[n] == 1) gx = gx + t;
        if (S::ex[n] == -1) gx = gx - t;
        if (S::ey[n] == 1) gy = gy + t;
        if (S::ey[n] == -1) gy = gy - t;
        if (S::ez[n] == 1) gz = gz + t;
        if (S::ez[n] == -1) gz = gz - t;
    }
    const double mg = -G;
    const double Fx = (mg * pc) * gx;
    const double Fy = (mg * pc) * gy;
    const double Fz = (mg * pc) * gz;

    const double ueq = u + (tau * Fx) / rho;
    const double veq = v + (tau * Fy) / rho;
    double weq, usq;
    if constexpr (S::D == 3) {
        weq = w + (tau * Fz) / rho;
        usq = (ueq * ueq + veq * veq) + weq * weq;
    } else {
        weq = 0.0;
        usq = ueq * ueq + veq * veq;
    }
    const double usq15 = 1.5 * usq;
    static_for<Q>([&](auto I) {
        constexpr int n = decltype(I)::value;
        const double ev = edot<S::ex[n], S::ey[n], S::ez[n]>(ueq, veq, weq);
        const double feq = (((1.0 + 3.0 * ev) + (4.5 * ev) * ev) - usq15) * (rho * S::w[n]);
        double out;
        if constexpr (MODE == 2) out = feq;
        else if constexpr (TAU1) out = f[n] - (f[n] - feq);
        else out = f[n] - (f[n] - feq) / tau;
        Pout[n * g.ps + c] = out;
    });
}

template <int Q>
__global__ void __launch_bounds__(kBlock)
    k_force(const double* __restrict__ psi, Grid g, double G, double* __restrict__ fx,
            double* __restrict__ fy, double* __restrict__ fz)
{
    using S = St<Q>;
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (c >= g.n) return;
    const Coord p = decompose(static_cast<uint32_t>(c), g);
    const Nbr nb = neighbours(p, g);
    const double pc = psi[c];
    double gx = 0.0, gy = 0.0, gz = 0.0;
#pragma unroll
    for (int n = 1; n < Q; n++) {
        const double t = S::w[n] * psi[at(nb, S::ex[n], S::ey[n], S::ez[n])];
        if (S::ex[n] == 1) gx = gx + t;
        if (S::ex[n] == -1) gx = gx - t;
        if (S::ey[n] == 1) gy = gy + t;
        if (S::ey[n] == -1) gy = gy - t;
        if (S::ez[n] == 1) gz = gz + t;
        if (S::ez[n] == -1) gz = gz - t;
    }
    const double mg = -G;
    fx[c] = (mg * pc) * gx;
    fy[c] = (mg * pc) * gy;
    fz[c] = (S::D == 3) ? (mg * pc) * gz : 0.0;
}

template <int Q, bool WALLS>
__global__ void __launch_bounds__(kBlock)
    k_moments(const double* __restrict__ P, int kind, Grid g, double* __restrict__ rho,
              double* __restrict__ u, double* __restrict__ v, double* __restrict__ w)
{
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (c >= g.n) return;
    const Coord p = decompose(static_cast<uint32_t>(c), g);
    const Nbr nb = neighbours(p, g);
    double f[Q];
    load_f<Q, WALLS>(P, g, kind, p, nb, c, f);
    double r, sx, sy, sz;
    moments<Q>(f, r, sx, sy, sz);
    rho[c] = r;
    u[c] = sx / r;
    v[c] = sy / r;
    w[c] = (St<Q>::D == 3) ? sz / r : 0.0;
}

template <int Q, bool WALLS>
__global__ void __launch_bounds__(kBlock)
    k_get_f(const double* __restrict__ P, int kind, Grid g, double* __restrict__ aos)
{
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (c >= g.n) return;
    const Coord p = decompose(static_cast<uint32_t>(c), g);
    const Nbr nb = neighbours(p, g);
    double f[Q];
    load_f<Q, WALLS>(P, g, kind, p, nb, c, f);
#pragma unroll
    for (int n = 0; n < Q; n++) aos[c * Q + n] = f[n];
}

template <int Q>
__global__ void __launch_bounds__(kBlock)
    k_set_f(const double* __restrict__ aos, Grid g, double* __restrict__ P)
{
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (c >= g.n) return;
#pragma unroll
    for (int n = 0; n < Q; n++) P[n * g.ps + c] = aos[c * Q + n];
}

// Fixed-shape block reduction (kConsBlock x kConsPerThread nodes per
// block, lbm_internal.h): deterministic for a given lattice.

template <int Q, bool WALLS>
__global__ void __launch_bounds__(kConsBlock)
    k_conservation(const double* __restrict__ P, int kind, Grid g, double* __restrict__ part)
{
    using S = St<Q>;
    __shared__ double red[4][kConsBlock];
    double m = 0.0, jx = 0.0, jy = 0.0, jz = 0.0;
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kConsBlock * kConsPerThread;
    for (int i = 0; i < kConsPerThread; i++) {
        const int64_t c = base + static_cast<int64_t>(i) * kConsBlock + threadIdx.x;
        if (c >= g.n) break;
        const Coord p = decompose(static_cast<uint32_t>(c), g);
        const Nbr nb = neighbours(p, g);
        double f[Q];
        load_f<Q, WALLS>(P, g, kind, p, nb, c, f);
        double r, sx, sy, sz;
        moments<Q>(f, r, sx, sy, sz);
        m += r; jx += sx; jy += sy; jz += sz;
        (void)S::D;
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

// ---- two components (mcmp2D / mcmp3D) ---------------------------------------
//
// Components A and B share one population buffer: component c holds planes
// c*Q .. c*Q+Q-1.  One node's step pulls both, forms the cross forces from the
// other component's psi at the neighbours and the common velocity, and
// collides each component with it, in the reference's operation order:
//   forces    mcmp2D.cpp:153-182 / mcmp3D.cpp:163-213
//             F_A = ((-G)*psi_A(x)) * sum (w*e)*psi_B(x+e), F_B likewise
//   common u  mcmp2D.cpp:185-196 with LBfluid2D.cpp:179-192
//             u' = ((u_A*r_A)/tau_A + (u_B*r_B)/tau_B) / (r_A/tau_A + r_B/tau_B)
//   collide   LBfluid2D.cpp:235-249 / LBfluid3D.cpp:259-276 with u = u'
// MODE 0: moments of the pulled f (mcmp3D's macros at the start of every
// step; mcmp2D's arrays after its first step); 1: rho = rho0 and u = 0
// (mcmp2D's first step: the arrays hold the initial densities and the
// common velocity initLatticeBoltzmann left, which is 0); 2: the initial
// equilibrium (mcmp2D.cpp:83-87, mcmp3D.cpp:87-91).

// sum_n (w_n e_n) psi(x + e_n) in stencil order (zero terms skipped: exact).
// FILM: neighbours in the wall rows (first and last row of the last axis)
// contribute psi_sol instead (mcmp2DFilm.cpp:173-181, mcmp3DFilm.cpp:199-208);
// last is this node's coordinate on that axis, nl its extent.
template <int Q, bool FILM = false>
__device__ __forceinline__ void psi_sums(const double* __restrict__ psi, const Nbr& nb, double& gx,
                                         double& gy, double& gz, int last = 0, int nl = 0,
                                         double psi_sol = 0.0)
{
    using S = St<Q>;
    gx = 0.0; gy = 0.0; gz = 0.0;
#pragma unroll
    for (int n = 1; n < Q; n++) {
        double pn;
        if constexpr (FILM) {
            const int ln = last + ((S::D == 3) ? S::ez[n] : S::ey[n]);
            pn = (ln == 0 || ln == nl - 1) ? psi_sol : psi[at(nb, S::ex[n], S::ey[n], S::ez[n])];
        } else {
            pn = psi[at(nb, S::ex[n], S::ey[n], S::ez[n])];
        }
        const double t = S::w[n] * pn;
        if (S::ex[n] == 1) gx = gx + t;
        if (S::ex[n] == -1) gx = gx - t;
        if (S::ey[n] == 1) gy = gy + t;
        if (S::ey[n] == -1) gy = gy - t;
        if (S::ez[n] == 1) gz = gz + t;
        if (S::ez[n] == -1) gz = gz - t;
    }
}

// BGK collision of one component with velocity u' + tau*F/rho; writes the
// post-collision values (MODE 2: the equilibrium itself).
template <int Q, int MODE>
__device__ __forceinline__ void collide_one(const double (&f)[Q], double rho, double up, double vp,
                                            double wp, double Fx, double Fy, double Fz, double tau,
                                            double* __restrict__ P, int64_t ps, int64_t c)
{
    using S = St<Q>;
    const double ueq = up + (tau * Fx) / rho;
    const double veq = vp + (tau * Fy) / rho;
    double weq, usq;
    if constexpr (S::D == 3) {
        weq = wp + (tau * Fz) / rho;
        usq = (ueq * ueq + veq * veq) + weq * weq;
    } else {
        weq = 0.0;
        usq = ueq * ueq + veq * veq;
    }
    const double usq15 = 1.5 * usq;
    static_for<Q>([&](auto I) {
        constexpr int n = decltype(I)::value;
        const double ev = edot<S::ex[n], S::ey[n], S::ez[n]>(ueq, veq, weq);
        const double feq = (((1.0 + 3.0 * ev) + (4.5 * ev) * ev) - usq15) * (rho * S::w[n]);
        P[n * ps + c] = (MODE == 2) ? feq : f[n] - (f[n] - feq) / tau;
    });
}

// Film parameters (mcmp2DFilm / mcmp3DFilm): the wall densities of A and B
// ("LBApp/rhoA_sol", "rhoB_sol"); their psi is formed in the kernel, with the
// same exp as every other psi.
struct Film {
    double rhoA, rhoB;
};

template <int Q, int MODE, bool FILM>
__global__ void __launch_bounds__(kBlock)
    k_collide2(const double* __restrict__ Pin, int kind, double* __restrict__ Pout,
               const double* __restrict__ psiA, const double* __restrict__ psiB,
               const double* __restrict__ rho0A, const double* __restrict__ rho0B, Grid g,
               double tauA, double tauB, double G, Film film)
{
    using S = St<Q>;
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (c >= g.n) return;
    const Coord p = decompose(static_cast<uint32_t>(c), g);
    const Nbr nb = neighbours(p, g);
    const int64_t off = static_cast<int64_t>(Q) * g.ps;   // component B's planes
    // film: the first and last row of the last axis are walls
    const int last = (S::D == 3) ? p.z : p.y;
    const int nl = (S::D == 3) ? g.nz : g.ny;
    const bool wall_row = FILM && (last == 0 || last == nl - 1);

    double fA[Q], fB[Q];
    double rA, rB, uA, vA, wA, uB, vB, wB;
    if constexpr (MODE == 0) {
        load_f<Q, FILM>(Pin, g, kind, p, nb, c, fA);
        load_f<Q, FILM>(Pin + off, g, kind, p, nb, c, fB);
        double sx, sy, sz;
        moments<Q>(fA, rA, sx, sy, sz);
        uA = sx / rA; vA = sy / rA; wA = sz / rA;
        moments<Q>(fB, rB, sx, sy, sz);
        uB = sx / rB; vB = sy / rB; wB = sz / rB;
    } else {
        if constexpr (MODE == 1) {
            load_f<Q, FILM>(Pin, g, kind, p, nb, c, fA);
            load_f<Q, FILM>(Pin + off, g, kind, p, nb, c, fB);
        }
        rA = rho0A[c]; rB = rho0B[c];
        uA = vA = wA = uB = vB = wB = 0.0;
    }

    if (wall_row) {
        // wall rows: no force and no common velocity -- the film models'
        // force loop skips them, so their force arrays keep the zeros they
        // were allocated with and their u arrays the component's own
        // velocity (mcmp2DFilm.cpp:153, mcmp3DFilm.cpp:176)
        collide_one<Q, MODE>(fA, rA, uA, vA, wA, 0.0, 0.0, 0.0, tauA, Pout, g.ps, c);
        collide_one<Q, MODE>(fB, rB, uB, vB, wB, 0.0, 0.0, 0.0, tauB, Pout + off, g.ps, c);
        return;
    }
    double gAx, gAy, gAz, gBx, gBy, gBz;
    const double psolA = FILM ? psi_of(film.rhoA) : 0.0, psolB = FILM ? psi_of(film.rhoB) : 0.0;
    psi_sums<Q, FILM>(psiA, nb, gAx, gAy, gAz, last, nl, psolA);
    psi_sums<Q, FILM>(psiB, nb, gBx, gBy, gBz, last, nl, psolB);
    const double mg = -G;
    const double pa = mg * psiA[c], pb = mg * psiB[c];

    const double rdA = rA / tauA, rdB = rB / tauB;
    const double den = rdA + rdB;
    const double up = ((uA * rA) / tauA + (uB * rB) / tauB) / den;
    const double vp = ((vA * rA) / tauA + (vB * rB) / tauB) / den;
    const double wp = (S::D == 3) ? ((wA * rA) / tauA + (wB * rB) / tauB) / den : 0.0;

    collide_one<Q, MODE>(fA, rA, up, vp, wp, pa * gBx, pa * gBy, pa * gBz, tauA, Pout, g.ps, c);
    collide_one<Q, MODE>(fB, rB, up, vp, wp, pb * gAx, pb * gAy, pb * gAz, tauB, Pout + off, g.ps, c);
}

// Cross force on the component whose psi is psi_self: ((-G)*psi_self(x)) *
// sum (w*e)*psi_other(x+e) (the fx,fy,fz arrays of mcmp2D / mcmp3D).
template <int Q, bool FILM>
__global__ void __launch_bounds__(kBlock)
    k_force2(const double* __restrict__ psi_self, const double* __restrict__ psi_other, Grid g, double G,
             double rho_sol_other, double* __restrict__ fx, double* __restrict__ fy,
             double* __restrict__ fz)
{
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (c >= g.n) return;
    const Coord p = decompose(static_cast<uint32_t>(c), g);
    const Nbr nb = neighbours(p, g);
    const int last = (St<Q>::D == 3) ? p.z : p.y;
    const int nl = (St<Q>::D == 3) ? g.nz : g.ny;
    if (FILM && (last == 0 || last == nl - 1)) {
        fx[c] = 0.0;   // the wall rows' force arrays are never set
        fy[c] = 0.0;
        fz[c] = 0.0;
        return;
    }
    double gx, gy, gz;
    psi_sums<Q, FILM>(psi_other, nb, gx, gy, gz, last, nl, FILM ? psi_of(rho_sol_other) : 0.0);
    const double pc = -G * psi_self[c];
    fx[c] = pc * gx;
    fy[c] = pc * gy;
    fz[c] = (St<Q>::D == 3) ? pc * gz : 0.0;
}

// Common velocity from both components' moments, written over A's velocity
// arrays (the u arrays mcmp3D leaves after a step).  film_nz > 0: the first
// and last z rows are walls, where the arrays keep the velocity of
// component `own` (0: A's, already in place; 1: B's).
__global__ void __launch_bounds__(kBlock)
    k_common_velocity(int64_t n, int three_d, const double* __restrict__ rA, double* __restrict__ uA,
                      double* __restrict__ vA, double* __restrict__ wA, const double* __restrict__ rB,
                      const double* __restrict__ uB, const double* __restrict__ vB,
                      const double* __restrict__ wB, double tauA, double tauB, int film_nz, int own)
{
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (c >= n) return;
    if (film_nz > 0) {
        const int z = static_cast<int>(c % film_nz);
        if (z == 0 || z == film_nz - 1) {
            if (own == 1) {
                uA[c] = uB[c];
                vA[c] = vB[c];
                wA[c] = wB[c];
            }
            return;
        }
    }
    const double ra = rA[c], rb = rB[c];
    const double den = ra / tauA + rb / tauB;
    uA[c] = ((uA[c] * ra) / tauA + (uB[c] * rb) / tauB) / den;
    vA[c] = ((vA[c] * ra) / tauA + (vB[c] * rb) / tauB) / den;
    wA[c] = three_d ? ((wA[c] * ra) / tauA + (wB[c] * rb) / tauB) / den : 0.0;
}

inline unsigned blocks_for(int64_t n) { return static_cast<unsigned>((n + kBlock - 1) / kBlock); }


// ---- immersed-boundary capsule (capsule2D) ----------------------------------
// A D2Q9 fluid in the LBfluid2D step order whose force arrays hold a body
// force plus the membrane forces of the IB nodes spread with the bilinear
// kernel (capsule2D.cpp:123-150, IBcapsules2D.cpp:339-386), in place of the
// Shan-Chen force.  Each fluid node gathers the contributions of the IB nodes
// whose footprint covers it, in IB node order, which is the order the
// reference's scatter adds them in; the IB nodes of a block's x range are
// compacted into LDS in that order first.

// Footprint origin of an IB node, 0-based: the nodes it touches are
// (ox + a, oy + b), a, b in {0, 1} (IBcapsules2D.cpp:238-239 with i0 - 1).
// False where the reference skips the node (isMyNode, IBcapsules2D.cpp:
// 418-426: -1 <= x < nx, which excludes every non-finite x), and for y
// outside [-1, ny), where the reference would index outside its ghost shell.
__device__ __forceinline__ bool ib_origin(double x, double y, int nx, int ny, int& ox, int& oy)
{
    if (!(x >= -1.0 && x < static_cast<double>(nx))) return false;
    if (!(y >= -1.0 && y < static_cast<double>(ny))) return false;
    ox = static_cast<int>(floor(x));
    oy = static_cast<int>(floor(y));
    return true;
}

// Bilinear weight of lattice point (px, py) for a node at (x, y)
// (IBcapsules2D.cpp:376-378).
__device__ __forceinline__ double ib_weight(double x, double y, int px, int py)
{
    const double rx = x - static_cast<double>(px);
    const double ry = y - static_cast<double>(py);
    return (1.0 - fabs(rx)) * (1.0 - fabs(ry));
}

constexpr int kIbChunk = kBlock;

template <bool WALLS, int MODE>
__global__ void __launch_bounds__(kBlock)
    k_collide_ib(const double* __restrict__ Pin, int kind, double* __restrict__ Pout,
                 const double* __restrict__ rho0, Grid g, double tau, Capsule cap)
{
    __shared__ double s_x[kIbChunk], s_y[kIbChunk], s_fx[kIbChunk], s_fy[kIbChunk];
    __shared__ int s_ox[kIbChunk], s_oy[kIbChunk];
    __shared__ int s_count;
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    const bool live = c < g.n;
    const int64_t cl = live ? c : g.n - 1;
    const Coord p = decompose(static_cast<uint32_t>(cl), g);

    // force: body force on the rows between the walls, then the IB nodes
    // (capsule2D.cpp:123-139); none at the start (MODE 2: the zero forces of
    // capsule2D.cpp:104-105)
    double Fx = 0.0, Fy = 0.0;
    if constexpr (MODE != 2) {
        if (p.y >= 1 && p.y <= g.ny - 2) {
            Fx = cap.body_fx;
            Fy = cap.body_fy;
        }
        // x range of this block's nodes (rows of ny nodes, x slowest)
        const int64_t c_lo = static_cast<int64_t>(blockIdx.x) * kBlock;
        const int64_t c_hi = std::min<int64_t>(c_lo + kBlock, g.n) - 1;
        const int x_lo = static_cast<int>(c_lo / g.ny), x_hi = static_cast<int>(c_hi / g.ny);
        for (int base = 0; base < cap.nnodes; base += kIbChunk) {
            __syncthreads();
            if (threadIdx.x == 0) s_count = 0;
            __syncthreads();
            // compact the chunk's nodes that can touch x_lo..x_hi, in order:
            // one wavefront at a time, ballot + prefix within it
            const int n = base + static_cast<int>(threadIdx.x);
            int ox = 0, oy = 0;
            bool keep = false;
            double xn = 0.0, yn = 0.0;
            if (n < cap.nnodes) {
                xn = cap.x[n];
                yn = cap.y[n];
                keep = ib_origin(xn, yn, g.nx, g.ny, ox, oy) && ox + 1 >= x_lo && ox <= x_hi;
            }
            const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
            for (int wv = 0; wv < kBlock / 64; wv++) {
                if (wave == wv) {
                    const uint64_t m = __ballot(keep);
                    const int before = __popcll(m & ((1ull << lane) - 1ull));
                    const int slot = s_count + before;
                    if (keep) {
                        s_x[slot] = xn;
                        s_y[slot] = yn;
                        // membrane force, IBcapsules2D.cpp:480-483
                        s_fx[slot] = -cap.kstiff * (xn - cap.x0[n]);
                        s_fy[slot] = -cap.kstiff * (yn - cap.y0[n]);
                        s_ox[slot] = ox;
                        s_oy[slot] = oy;
                    }
                    __builtin_amdgcn_wave_barrier();
                    if (lane == 0) s_count += __popcll(m);
                }
                __syncthreads();
            }
            const int cnt = s_count;
            for (int k = 0; k < cnt; k++) {
                const int ax = p.x - s_ox[k], ay = p.y - s_oy[k];
                if ((ax == 0 || ax == 1) && (ay == 0 || ay == 1)) {
                    const double del = ib_weight(s_x[k], s_y[k], p.x, p.y);
                    Fx = Fx + del * s_fx[k];
                    Fy = Fy + del * s_fy[k];
                }
            }
        }
    }
    if (!live) return;
    if constexpr (MODE != 2) {
        cap.fx[c] = Fx;
        cap.fy[c] = Fy;
    }

    double f[9];
    double rho, u, v;
    if constexpr (MODE == 0) {
        const Nbr nb = neighbours(p, g);
        load_f<9, WALLS>(Pin, g, kind, p, nb, c, f);
        double sx, sy, sz;
        moments<9>(f, rho, sx, sy, sz);
        u = sx / rho;
        v = sy / rho;
    } else {
        if constexpr (MODE == 1) {
            const Nbr nb = neighbours(p, g);
            load_f<9, WALLS>(Pin, g, kind, p, nb, c, f);
        } else {
#pragma unroll
            for (int n = 0; n < 9; n++) f[n] = 0.0;
        }
        rho = rho0[c];
        u = 0.0;
        v = 0.0;
    }
    collide_one<9, MODE>(f, rho, u, v, 0.0, Fx, Fy, 0.0, tau, Pout, g.ps, c);
}

// After the fluid step (one workgroup): the force that the spread put on the
// ghost rows below and above the lattice, which LBfluid2D::zeroForces never
// clears (LBfluid2D.cpp:137-147), then per IB node the interpolated velocity
// u* = u + F/(2 rho) of the new fluid state (IBcapsules2D.cpp:218-287,
// LBfluid2D.cpp:164-172; the ghost rows read r = u = 0) and the move
// x += v dt (IBcapsules2D.cpp:463-469).
//
// The nodes are taken kIbStage at a time: their positions, footprint
// origins and membrane forces are staged in LDS for the ghost-row sums, and
// each of their footprint points is one thread's work item (pull the 9
// values, moments, the weighted u*), so the four points of a node are
// evaluated at once instead of one after another.  A node's velocity then
// adds its points' terms in the order of the reference's loops (px outer,
// py inner), so the sums round as they do there.
constexpr int kIbStage = 256;

template <bool WALLS>
__global__ void __launch_bounds__(1024) k_ib_finish(const double* __restrict__ P, int kind, Grid g, Capsule cap)
{
    static_assert(4 * kIbStage <= 1024, "one thread per footprint point of a staged node");
    __shared__ double s_x[kIbStage], s_y[kIbStage], s_fx[kIbStage], s_fy[kIbStage];
    __shared__ int s_ox[kIbStage], s_oy[kIbStage];
    __shared__ bool s_ok[kIbStage];
    __shared__ double s_u[4 * kIbStage], s_v[4 * kIbStage];
    __shared__ bool s_use[4 * kIbStage];
    const int nx = g.nx, ny = g.ny;
    const int tid = threadIdx.x, nt = blockDim.x;
    auto stage = [&](int base) {
        const int cnt = min(kIbStage, cap.nnodes - base);
        for (int k = tid; k < cnt; k += nt) {
            const int n = base + k;
            const double xn = cap.x[n], yn = cap.y[n];
            int ox = 0, oy = 0;
            s_ok[k] = ib_origin(xn, yn, nx, ny, ox, oy);
            s_x[k] = xn;
            s_y[k] = yn;
            s_ox[k] = ox;
            s_oy[k] = oy;
            // membrane force, IBcapsules2D.cpp:480-483
            s_fx[k] = -cap.kstiff * (xn - cap.x0[n]);
            s_fy[k] = -cap.kstiff * (yn - cap.y0[n]);
        }
        return cnt;
    };
    // ghost rows: each position adds the spread of every node, in node order
    for (int t0 = 0; t0 < 2 * nx; t0 += nt) {
        const int t = t0 + tid;
        const bool mine = t < 2 * nx;
        const int side = mine ? t / nx : 0, px = mine ? t - side * nx : 0;
        const int py = side ? ny : -1;
        double gx = mine ? cap.gfx[t] : 0.0, gy = mine ? cap.gfy[t] : 0.0;
        for (int base = 0; base < cap.nnodes; base += kIbStage) {
            __syncthreads();
            const int cnt = stage(base);
            __syncthreads();
            if (!mine) continue;
            for (int k = 0; k < cnt; k++) {
                if (!s_ok[k]) continue;
                const int ax = px - s_ox[k], ay = py - s_oy[k];
                if ((ax == 0 || ax == 1) && (ay == 0 || ay == 1)) {
                    const double del = ib_weight(s_x[k], s_y[k], px, py);
                    gx = gx + del * s_fx[k];
                    gy = gy + del * s_fy[k];
                }
            }
        }
        if (mine) {
            cap.gfx[t] = gx;
            cap.gfy[t] = gy;
        }
    }
    for (int n = tid; n < cap.nnodes; n += nt) {
        cap.nfx[n] = -cap.kstiff * (cap.x[n] - cap.x0[n]);
        cap.nfy[n] = -cap.kstiff * (cap.y[n] - cap.y0[n]);
    }
    __syncthreads();
    for (int base = 0; base < cap.nnodes; base += kIbStage) {
        const int cnt = min(kIbStage, cap.nnodes - base);
        // item i: point (i & 3) of node base + i / 4, point = 2 (px - ox) + (py - oy)
        if (tid < 4 * cnt) {
            const int k = tid >> 2, pt = tid & 3;
            const double xn = cap.x[base + k], yn = cap.y[base + k];
            int ox = 0, oy = 0;
            const bool ok = ib_origin(xn, yn, nx, ny, ox, oy);
            const int px = ox + (pt >> 1), py = oy + (pt & 1);
            const bool use = ok && px >= 0 && px <= nx - 1;
            double cu = 0.0, cv = 0.0;
            if (use) {
                const double del = ib_weight(xn, yn, px, py);
                double us, vs;
                if (py >= 0 && py < ny) {
                    const int64_t c = static_cast<int64_t>(px) * ny + py;
                    const Coord q{px, py, 0};
                    const Nbr nb = neighbours(q, g);
                    double f[9];
                    load_f<9, WALLS>(P, g, kind, q, nb, c, f);
                    double r, sx, sy, sz;
                    moments<9>(f, r, sx, sy, sz);
                    us = sx / r + (0.5 * cap.fx[c]) / r;
                    vs = sy / r + (0.5 * cap.fy[c]) / r;
                } else {
                    const int t = (py < 0 ? 0 : nx) + px;
                    us = 0.0 + (0.5 * cap.gfx[t]) / 0.0;
                    vs = 0.0 + (0.5 * cap.gfy[t]) / 0.0;
                }
                cu = del * us;
                cv = del * vs;
            }
            s_u[tid] = cu;
            s_v[tid] = cv;
            s_use[tid] = use;
        }
        __syncthreads();
        if (tid < cnt) {
            const int n = base + tid;
            double vx = 0.0, vy = 0.0;
            for (int pt = 0; pt < 4; pt++) {
                if (!s_use[4 * tid + pt]) continue;
                vx = vx + s_u[4 * tid + pt];
                vy = vy + s_v[4 * tid + pt];
            }
            const double xn = cap.x[n], yn = cap.y[n];
            cap.vx[n] = vx;
            cap.vy[n] = vy;
            cap.x[n] = xn + vx * cap.dt;
            cap.y[n] = yn + vy * cap.dt;
        }
        __syncthreads();
    }
}

}  // namespace

// ---- launchers -------------------------------------------------------------

hipError_t fp64_psi_from_rho(const double* rho, double* psi, int64_t n, hipStream_t s)
{
    k_psi_rho<<<blocks_for(n), kBlock, 0, s>>>(rho, psi, n);
    return hipGetLastError();
}

#define LBM_DISPATCH_QW(a, CALL)                                   \
    do {                                                          \
        if ((a).q == 9) {                                         \
            if ((a).walls) { constexpr int Q = 9; constexpr bool W = true; CALL; } \
            else { constexpr int Q = 9; constexpr bool W = false; CALL; }          \
        } else {                                                  \
            if ((a).walls) { constexpr int Q = 19; constexpr bool W = true; CALL; } \
            else { constexpr int Q = 19; constexpr bool W = false; CALL; }          \
        }                                                         \
    } while (0)

hipError_t fp64_psi_from_pops(const StepArgs& a, const Grid& g, const double* P, int kind,
                              double* psi, hipStream_t s)
{
    LBM_DISPATCH_QW(a, (k_psi_pops<Q, W><<<blocks_for(g.n), kBlock, 0, s>>>(P, kind, psi, g)));
    return hipGetLastError();
}

hipError_t fp64_collide(const StepArgs& a, const Grid& g, const double* Pin, int kind,
                        double* Pout, const double* psi, const double* rho0, int mode,
                        hipStream_t s)
{
    const unsigned nb = blocks_for(g.n);
    const bool t1 = a.tau == 1.0;
    if (mode == 0 && t1)
        LBM_DISPATCH_QW(a, (k_collide<Q, W, 0, true><<<nb, kBlock, 0, s>>>(Pin, kind, Pout, psi, rho0, g, a.tau, a.G)));
    else if (mode == 0)
        LBM_DISPATCH_QW(a, (k_collide<Q, W, 0><<<nb, kBlock, 0, s>>>(Pin, kind, Pout, psi, rho0, g, a.tau, a.G)));
    else if (mode == 1 && t1)
        LBM_DISPATCH_QW(a, (k_collide<Q, W, 1, true><<<nb, kBlock, 0, s>>>(Pin, kind, Pout, psi, rho0, g, a.tau, a.G)));
    else if (mode == 1)
        LBM_DISPATCH_QW(a, (k_collide<Q, W, 1><<<nb, kBlock, 0, s>>>(Pin, kind, Pout, psi, rho0, g, a.tau, a.G)));
    else
        LBM_DISPATCH_QW(a, (k_collide<Q, W, 2><<<nb, kBlock, 0, s>>>(Pin, kind, Pout, psi, rho0, g, a.tau, a.G)));
    return hipGetLastError();
}

hipError_t fp64_force(const StepArgs& a, const Grid& g, const double* psi, double* fx, double* fy,
                      double* fz, hipStream_t s)
{
    if (a.q == 9) k_force<9><<<blocks_for(g.n), kBlock, 0, s>>>(psi, g, a.G, fx, fy, fz);
    else k_force<19><<<blocks_for(g.n), kBlock, 0, s>>>(psi, g, a.G, fx, fy, fz);
    return hipGetLastError();
}

hipError_t fp64_moments(const StepArgs& a, const Grid& g, const double* P, int kind, double* rho,
                        double* u, double* v, double* w, hipStream_t s)
{
    LBM_DISPATCH_QW(a, (k_moments<Q, W><<<blocks_for(g.n), kBlock, 0, s>>>(P, kind, g, rho, u, v, w)));
    return hipGetLastError();
}

hipError_t fp64_get_f(const StepArgs& a, const Grid& g, const double* P, int kind, double* aos,
                      hipStream_t s)
{
    LBM_DISPATCH_QW(a, (k_get_f<Q, W><<<blocks_for(g.n), kBlock, 0, s>>>(P, kind, g, aos)));
    return hipGetLastError();
}

hipError_t fp64_set_f(const StepArgs& a, const Grid& g, const double* aos, double* P,
                      hipStream_t s)
{
    if (a.q == 9) k_set_f<9><<<blocks_for(g.n), kBlock, 0, s>>>(aos, g, P);
    else k_set_f<19><<<blocks_for(g.n), kBlock, 0, s>>>(aos, g, P);
    return hipGetLastError();
}

hipError_t fp64_collide2(const StepArgs& a, const Grid& g, const double* Pin, int kind, double* Pout,
                         const double* psi, const double* rho0, int mode, hipStream_t s)
{
    const unsigned nb = blocks_for(g.n);
    const double* psiB = psi + g.n;
    const double* r0B = rho0 + g.n;
    const Film film{a.rho_sol_a, a.rho_sol_b};
    auto go = [&](auto qtag, auto ftag) {
        constexpr int Q = decltype(qtag)::value;
        constexpr bool F = decltype(ftag)::value;
        if (mode == 0)
            k_collide2<Q, 0, F><<<nb, kBlock, 0, s>>>(Pin, kind, Pout, psi, psiB, rho0, r0B, g, a.tau, a.tau_b,
                                                       a.G, film);
        else if (mode == 1)
            k_collide2<Q, 1, F><<<nb, kBlock, 0, s>>>(Pin, kind, Pout, psi, psiB, rho0, r0B, g, a.tau, a.tau_b,
                                                       a.G, film);
        else
            k_collide2<Q, 2, F><<<nb, kBlock, 0, s>>>(Pin, kind, Pout, psi, psiB, rho0, r0B, g, a.tau, a.tau_b,
                                                       a.G, film);
    };
    using T9 = std::integral_constant<int, 9>;
    using T19 = std::integral_constant<int, 19>;
    if (a.q == 9) {
        if (a.walls) go(T9{}, std::true_type{});
        else go(T9{}, std::false_type{});
    } else {
        if (a.walls) go(T19{}, std::true_type{});
        else go(T19{}, std::false_type{});
    }
    return hipGetLastError();
}

hipError_t fp64_force2(const StepArgs& a, const Grid& g, const double* psi_self, const double* psi_other,
                       int other, double* fx, double* fy, double* fz, hipStream_t s)
{
    const double ps = other == 0 ? a.rho_sol_a : a.rho_sol_b;
    const unsigned nb = blocks_for(g.n);
    if (a.q == 9) {
        if (a.walls) k_force2<9, true><<<nb, kBlock, 0, s>>>(psi_self, psi_other, g, a.G, ps, fx, fy, fz);
        else k_force2<9, false><<<nb, kBlock, 0, s>>>(psi_self, psi_other, g, a.G, ps, fx, fy, fz);
    } else {
        if (a.walls) k_force2<19, true><<<nb, kBlock, 0, s>>>(psi_self, psi_other, g, a.G, ps, fx, fy, fz);
        else k_force2<19, false><<<nb, kBlock, 0, s>>>(psi_self, psi_other, g, a.G, ps, fx, fy, fz);
    }
    return hipGetLastError();
}

hipError_t fp64_common_velocity(const StepArgs& a, const Grid& g, const double* rA, double* uA, double* vA,
                                double* wA, const double* rB, const double* uB, const double* vB,
                                const double* wB, int own, hipStream_t s)
{
    const int film_nz = (a.walls && a.q == 19) ? g.nz : 0;
    k_common_velocity<<<blocks_for(g.n), kBlock, 0, s>>>(g.n, a.q == 19 ? 1 : 0, rA, uA, vA, wA, rB, uB, vB,
                                                         wB, a.tau, a.tau_b, film_nz, own);
    return hipGetLastError();
}

int fp64_conservation_blocks(const Grid& g) { return static_cast<int>(conservation_blocks(g.n)); }

hipError_t fp64_conservation(const StepArgs& a, const Grid& g, const double* P, int kind,
                             double* partials, int* nblocks, hipStream_t s)
{
    const int nb = fp64_conservation_blocks(g);
    *nblocks = nb;
    LBM_DISPATCH_QW(a, (k_conservation<Q, W><<<nb, kConsBlock, 0, s>>>(P, kind, g, partials)));
    return hipGetLastError();
}


hipError_t fp64_capsule_collide(const StepArgs& a, const Grid& g, const double* Pin, int kind, double* Pout,
                                const double* rho0, int mode, const Capsule& cap, hipStream_t s)
{
    if (a.q != 9) return hipErrorInvalidValue;
    const unsigned nb = static_cast<unsigned>(blocks_for(g.n));
    auto go = [&](auto walls) {
        constexpr bool W = decltype(walls)::value;
        if (mode == 0) k_collide_ib<W, 0><<<nb, kBlock, 0, s>>>(Pin, kind, Pout, rho0, g, a.tau, cap);
        else if (mode == 1) k_collide_ib<W, 1><<<nb, kBlock, 0, s>>>(Pin, kind, Pout, rho0, g, a.tau, cap);
        else k_collide_ib<W, 2><<<nb, kBlock, 0, s>>>(Pin, kind, Pout, rho0, g, a.tau, cap);
    };
    if (a.walls) go(std::true_type{});
    else go(std::false_type{});
    return hipGetLastError();
}

hipError_t fp64_capsule_finish(const StepArgs& a, const Grid& g, const double* P, int kind, const Capsule& cap,
                               hipStream_t s)
{
    if (a.q != 9) return hipErrorInvalidValue;
    if (a.walls) k_ib_finish<true><<<1, 1024, 0, s>>>(P, kind, g, cap);
    else k_ib_finish<false><<<1, 1024, 0, s>>>(P, kind, g, cap);
    return hipGetLastError();
}

}  // namespace lbm

// ---- test hook (include/mesobasic_lbm_testing.h) ---------------------------

namespace {
__global__ void k_probe_exp(const double* __restrict__ x, double* __restrict__ y, int64_t n)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (i < n) y[i] = lbm::exp_libm(x[i]);
}
}  // namespace

extern "C" int lbm_probe_exp(const double* x, double* y, int64_t n)
{
    if (n <= 0) return 0;
    double* d = nullptr;
    if (hipMalloc(&d, 2 * n * sizeof(double)) != hipSuccess) return -2;
    int rc = 0;
    if (hipMemcpy(d, x, n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) rc = -3;
    if (!rc) {
        k_probe_exp<<<static_cast<unsigned>((n + 255) / 256), 256>>>(d, d + n, n);
        if (hipGetLastError() != hipSuccess) rc = -3;
    }
    if (!rc && hipMemcpy(y, d + n, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) rc = -3;
    (void)hipFree(d);
    return rc;
}

// Streaming reference kernels over `bytes` (a multiple of 16): the points
// that calibrate HBM traffic counters (FETCH_SIZE / WRITE_SIZE) and bound
// the achievable bandwidth.  mode 0: copy, 4 x 16 B per thread; 1: copy,
// 8 x 16 B per thread; 2: as 1 with non-temporal loads and stores;
// 3: read-only (sum of 8 x 16 B per thread into one word per block);
// 4: copy, 1 x 16 B per thread (the fastest copy measured on MI355X).
namespace {
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

template <int K, bool NT>
__global__ void __launch_bounds__(256) k_probe_copy(const u32x4* __restrict__ src,
                                                    u32x4* __restrict__ dst, int64_t n16)
{
    const int64_t base = static_cast<int64_t>(blockIdx.x) * (256 * K) + threadIdx.x;
    u32x4 v[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int64_t i = base + k * 256;
        if (i < n16) {
            if constexpr (NT) v[k] = __builtin_nontemporal_load(src + i);
            else v[k] = src[i];
        }
    }
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int64_t i = base + k * 256;
        if (i < n16) {
            if constexpr (NT) __builtin_nontemporal_store(v[k], dst + i);
            else dst[i] = v[k];
        }
    }
}

__global__ void __launch_bounds__(256) k_probe_read(const u32x4* __restrict__ src,
                                                    uint32_t* __restrict__ out, int64_t n16)
{
    const int64_t base = static_cast<int64_t>(blockIdx.x) * (256 * 8) + threadIdx.x;
    uint32_t acc = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int64_t i = base + k * 256;
        if (i < n16) {
            const u32x4 v = src[i];
            acc ^= v.x ^ v.y ^ v.z ^ v.w;
        }
    }
    // one word per block (so the loads are not dead); lane 0 of each wave
    if ((threadIdx.x & 63) == 0) out[blockIdx.x * 4 + (threadIdx.x >> 6)] = acc;
}
}  // namespace

extern "C" int lbm_probe_copy(int64_t bytes, int reps, int mode, double* ms)
{
    if (bytes <= 0 || bytes % 16 != 0 || reps < 1 || mode < 0 || mode > 4) return -1;
    void* a = nullptr;
    void* b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess) return -2;
    if (hipMalloc(&b, bytes) != hipSuccess) { (void)hipFree(a); return -2; }
    int rc = 0;
    if (hipMemset(a, 1, bytes) != hipSuccess) rc = -3;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    const int64_t n16 = bytes / 16;
    const int per = (mode == 0) ? 4 : (mode == 4 ? 1 : 8);
    const unsigned grid = static_cast<unsigned>((n16 + 256 * per - 1) / (256 * per));
    // the read-only kernel writes 4 words per block: b has room (grid*16 <= bytes)
    if (!rc) {
        // one untimed run first (r = -1): the timed runs see warm TLBs and clocks
        for (int r = -1; r < reps; r++) {
            if (r == 0) (void)hipEventRecord(e0, nullptr);
            const u32x4* src = static_cast<const u32x4*>(a);
            u32x4* dst = static_cast<u32x4*>(b);
            if (mode == 0) k_probe_copy<4, false><<<grid, 256>>>(src, dst, n16);
            else if (mode == 1) k_probe_copy<8, false><<<grid, 256>>>(src, dst, n16);
            else if (mode == 2) k_probe_copy<8, true><<<grid, 256>>>(src, dst, n16);
            else if (mode == 4) k_probe_copy<1, false><<<grid, 256>>>(src, dst, n16);
            else k_probe_read<<<grid, 256>>>(src, static_cast<uint32_t*>(b), n16);
        }
        (void)hipEventRecord(e1, nullptr);
        if (hipEventSynchronize(e1) != hipSuccess) rc = -3;
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

// Device-to-device hipMemcpyAsync of `bytes` between two buffers it
// allocates, `reps` times after one untimed run; *ms = average time per
// copy (HIP events).  The runtime's own copy path: an attainable-bandwidth
// figure independent of the probe kernels above.
extern "C" int lbm_probe_memcpy(int64_t bytes, int reps, double* ms)
{
    if (bytes <= 0 || reps < 1) return -1;
    void* a = nullptr;
    void* b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess) return -2;
    if (hipMalloc(&b, bytes) != hipSuccess) { (void)hipFree(a); return -2; }
    int rc = 0;
    if (hipMemset(a, 1, bytes) != hipSuccess) rc = -3;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    if (!rc) {
        for (int r = -1; r < reps && !rc; r++) {
            if (r == 0) (void)hipEventRecord(e0, nullptr);
            if (hipMemcpyAsync(b, a, bytes, hipMemcpyDeviceToDevice, nullptr) != hipSuccess) rc = -3;
        }
        (void)hipEventRecord(e1, nullptr);
        if (hipEventSynchronize(e1) != hipSuccess) rc = -3;
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
