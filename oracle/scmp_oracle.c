/*
 * scmp_oracle.c -- CPU restatement of the reference SCMP lattice-Boltzmann
 * step.  TEST INFRASTRUCTURE ONLY: the HIP product path never links or calls
 * this file; tests, the smoke check and the bench's CPU-baseline leg do.
 *
 * Every loop below keeps the reference's array layout (AoS, one-node ghost
 * shell, x slowest / last axis fastest), loop order and left-to-right
 * floating-point evaluation order, so that results are bit-identical to the
 * reference compiled with g++ -O3 on x86-64 (no FMA contraction; build this
 * file with -ffp-contract=off and no -march flags).  File:line citations are
 * to the reference tree (flowzario/mesoBasic).
 */
#include "scmp_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

struct scmp_oracle {
    int q, dim, order, walls, nthreads;
    int nx, ny, nz;                /* interior sizes (nz = 1 in 2-D)      */
    int64_t gx, gy, gz;            /* sizes with the ghost shell          */
    int64_t deli, delj, delk;      /* LBfluid3D.cpp:27-29 / LBfluid2D.cpp:52-53 */
    int64_t gn;                    /* number of nodes with ghosts          */
    double tau, G;
    int exi[19], eyi[19], ezi[19];
    double ex[19], ey[19], ez[19], wa[19];
    double *r, *u, *v, *w, *fx, *fy, *fz, *f, *fs;
    double (*expfn)(double);
};

/* Stencil::setStencil, Stencil.cpp:5-120.  The direction order fixes every
 * summation order below and the bounce-back pairs. */
static void set_stencil(scmp_oracle* h)
{
    static const int d2[9][2] = {{0, 0}, {1, 0}, {-1, 0}, {0, 1}, {0, -1},
                                 {1, 1}, {-1, -1}, {-1, 1}, {1, -1}};
    static const int d3[19][3] = {
        {0, 0, 0},  {1, 0, 0},  {-1, 0, 0}, {0, 1, 0},  {0, -1, 0},
        {0, 0, 1},  {0, 0, -1}, {1, 1, 0},  {-1, -1, 0}, {1, 0, 1},
        {-1, 0, -1}, {0, 1, 1}, {0, -1, -1}, {1, -1, 0}, {-1, 1, 0},
        {1, 0, -1}, {-1, 0, 1}, {0, 1, -1}, {0, -1, 1}};
    if (h->q == 9) {
        for (int n = 0; n < 9; n++) {
            h->exi[n] = d2[n][0]; h->eyi[n] = d2[n][1]; h->ezi[n] = 0;
        }
        h->wa[0] = 4.0 / 9.0;                                   /* :32 */
        for (int n = 1; n < 5; n++) h->wa[n] = 1.0 / 9.0;       /* :33-36 */
        for (int n = 5; n < 9; n++) h->wa[n] = 1.0 / 36.0;      /* :37-40 */
    } else {
        for (int n = 0; n < 19; n++) {
            h->exi[n] = d3[n][0]; h->eyi[n] = d3[n][1]; h->ezi[n] = d3[n][2];
        }
        h->wa[0] = 1.0 / 3.0;                                   /* :79 */
        for (int n = 1; n < 7; n++) h->wa[n] = 1.0 / 18.0;      /* :80-85 */
        for (int n = 7; n < 19; n++) h->wa[n] = 1.0 / 36.0;     /* :86-97 */
    }
    for (int n = 0; n < h->q; n++) {
        h->ex[n] = (double)h->exi[n];
        h->ey[n] = (double)h->eyi[n];
        h->ez[n] = (double)h->ezi[n];
    }
}

scmp_oracle* scmp_oracle_create(int q, int nx, int ny, int nz, double tau,
                                double G, int order, int walls, int nthreads)
{
    if (!(q == 9 || q == 19) || nx < 1 || ny < 1 || nz < 1) return NULL;
    if (q == 9 && nz != 1) return NULL;
    scmp_oracle* h = (scmp_oracle*)calloc(1, sizeof(*h));
    if (!h) return NULL;
    h->q = q;
    h->dim = (q == 9) ? 2 : 3;
    h->order = order;
    h->walls = walls;
    h->nthreads = nthreads < 1 ? 1 : nthreads;
    h->nx = nx; h->ny = ny; h->nz = nz;
    h->tau = tau; h->G = G;
    h->gx = nx + 2;
    h->gy = ny + 2;
    if (h->dim == 3) {
        h->gz = nz + 2;
        h->deli = h->gy * h->gz; h->delj = h->gz; h->delk = 1;
    } else {
        h->gz = 1;
        h->deli = h->gy; h->delj = 1; h->delk = 0;
    }
    h->gn = h->gx * h->gy * h->gz;
    set_stencil(h);
    h->expfn = exp;
    size_t sn = (size_t)h->gn * sizeof(double);
    size_t sf = (size_t)h->gn * (size_t)q * sizeof(double);
    h->r = (double*)calloc(1, sn);  h->u = (double*)calloc(1, sn);
    h->v = (double*)calloc(1, sn);  h->w = (double*)calloc(1, sn);
    h->fx = (double*)calloc(1, sn); h->fy = (double*)calloc(1, sn);
    h->fz = (double*)calloc(1, sn);
    h->f = (double*)calloc(1, sf);  h->fs = (double*)calloc(1, sf);
    if (!h->r || !h->u || !h->v || !h->w || !h->fx || !h->fy || !h->fz ||
        !h->f || !h->fs) {
        scmp_oracle_destroy(h);
        return NULL;
    }
    return h;
}

void scmp_oracle_destroy(scmp_oracle* h)
{
    if (!h) return;
    free(h->r); free(h->u); free(h->v); free(h->w);
    free(h->fx); free(h->fy); free(h->fz); free(h->f); free(h->fs);
    free(h);
}

int64_t scmp_oracle_nodes(const scmp_oracle* h)
{
    return (int64_t)h->nx * h->ny * h->nz;
}

void scmp_oracle_set_exp(scmp_oracle* h, double (*fn)(double))
{
    h->expfn = fn ? fn : exp;
}

static inline int64_t nd(const scmp_oracle* h, int64_t i, int64_t j, int64_t k)
{
    return i * h->deli + j * h->delj + k * h->delk;
}

/* scmp3D::psi, scmp3D.cpp:187-190 (rho0 = 1 is hard-wired there). */
static inline double psi(const scmp_oracle* h, double rho)
{
    return (1.0 - h->expfn(-rho));
}

/* Interior iteration: x slabs over threads, the rest serial. */
#define FOR_INTERIOR_BEGIN(h)                                                \
    _Pragma("omp parallel for num_threads(h->nthreads) schedule(static)")    \
    for (int i = 1; i < (h)->nx + 1; i++) {                                  \
        for (int j = 1; j < (h)->ny + 1; j++) {                              \
            for (int k = ((h)->dim == 3 ? 1 : 0);                            \
                 k < ((h)->dim == 3 ? (h)->nz + 1 : 1); k++) {               \
                int64_t ndx = nd((h), i, j, k);
#define FOR_INTERIOR_END }}}

/* Periodic fill of the ghost shell of a field with `per` values per node.
 * Order as LBfluid3D::ghostNodesRho / ghostNodesStreaming
 * (LBfluid3D.cpp:430-506): x planes first (the single-rank mpiExchange,
 * :515-543, with both neighbours equal to self), then y over all i, then z
 * over all i and j, so edges and corners come out periodic.
 * LBfluid2D.cpp:495-543 is the same without z. */
static void ghost_fill(const scmp_oracle* h, double* a, int per)
{
    const int64_t plane = h->deli * per;
    /* send-left/recv-right then send-right/recv-left (mpiExchange :524-541) */
    memcpy(a + (int64_t)(h->nx + 1) * plane, a + 1 * plane,
           (size_t)plane * sizeof(double));
    memcpy(a + 0, a + (int64_t)h->nx * plane, (size_t)plane * sizeof(double));
    for (int64_t i = 0; i < h->nx + 2; i++) {
        for (int64_t k = 0; k < h->gz; k++) {
            for (int n = 0; n < per; n++) {
                a[nd(h, i, 0, k) * per + n] = a[nd(h, i, h->ny, k) * per + n];
                a[nd(h, i, h->ny + 1, k) * per + n] = a[nd(h, i, 1, k) * per + n];
            }
        }
    }
    if (h->dim == 3) {
        for (int64_t i = 0; i < h->nx + 2; i++) {
            for (int64_t j = 0; j < h->ny + 2; j++) {
                for (int n = 0; n < per; n++) {
                    a[nd(h, i, j, 0) * per + n] = a[nd(h, i, j, h->nz) * per + n];
                    a[nd(h, i, j, h->nz + 1) * per + n] = a[nd(h, i, j, 1) * per + n];
                }
            }
        }
    }
}

/* scmp3D::calculateShanChenForces, scmp3D.cpp:141-179; in 2-D the same
 * formula on the D2Q9 stencil (the loop shape of mcmp2D.cpp:145-183 for a
 * single component). */
static void shan_chen_forces(scmp_oracle* h)
{
    const int q = h->q;
    FOR_INTERIOR_BEGIN(h)
        double rhoA = h->r[ndx];
        double psi_rhoA = psi(h, rhoA);
        double GsumxA = 0.0, GsumyA = 0.0, GsumzA = 0.0;
        for (int n = 0; n < q; n++) {
            int64_t nbr = nd(h, i + h->exi[n], j + h->eyi[n],
                             h->dim == 3 ? k + h->ezi[n] : 0);
            double psi_nbr = psi(h, h->r[nbr]);
            GsumxA += h->wa[n] * h->ex[n] * psi_nbr;
            GsumyA += h->wa[n] * h->ey[n] * psi_nbr;
            GsumzA += h->wa[n] * h->ez[n] * psi_nbr;
        }
        h->fx[ndx] = -h->G * psi_rhoA * GsumxA;
        h->fy[ndx] = -h->G * psi_rhoA * GsumyA;
        if (h->dim == 3) h->fz[ndx] = -h->G * psi_rhoA * GsumzA;
    FOR_INTERIOR_END
}

/* Equilibrium at a node with the velocity shift ueq = u + tau*F/rho.
 * LBfluid3D.cpp:180-200 (setFtoFeq) and :259-276 (collision);
 * LBfluid2D.cpp:201-218 and :235-249 for D2Q9 (no z terms). */
static inline void shifted_velocity(const scmp_oracle* h, int64_t ndx,
                                    double* ueq, double* veq, double* weq,
                                    double* usq)
{
    const double tau = h->tau;
    *ueq = h->u[ndx] + tau * h->fx[ndx] / h->r[ndx];
    *veq = h->v[ndx] + tau * h->fy[ndx] / h->r[ndx];
    if (h->dim == 3) {
        *weq = h->w[ndx] + tau * h->fz[ndx] / h->r[ndx];
        *usq = *ueq * *ueq + *veq * *veq + *weq * *weq;
    } else {
        *weq = 0.0;
        *usq = *ueq * *ueq + *veq * *veq;
    }
}

static inline double feq_n(const scmp_oracle* h, int n, double rho,
                           double ueq, double veq, double weq, double usq)
{
    double evel = (h->dim == 3)
                      ? h->ex[n] * ueq + h->ey[n] * veq + h->ez[n] * weq
                      : h->ex[n] * ueq + h->ey[n] * veq;
    double feq = 1.0 + 3.0 * evel + 4.5 * evel * evel - 1.5 * usq;
    feq *= rho * h->wa[n];
    return feq;
}

static void set_f_to_feq(scmp_oracle* h)
{
    const int q = h->q;
    FOR_INTERIOR_BEGIN(h)
        double ueq, veq, weq, usq;
        shifted_velocity(h, ndx, &ueq, &veq, &weq, &usq);
        for (int n = 0; n < q; n++)
            h->f[ndx * q + n] = feq_n(h, n, h->r[ndx], ueq, veq, weq, usq);
    FOR_INTERIOR_END
}

/* LBfluid3D::macros, LBfluid3D.cpp:208-244 (bare velocity, no F/2 shift);
 * LBfluid2D.cpp:285-301 for D2Q9. */
static void macros(scmp_oracle* h)
{
    const int q = h->q;
    FOR_INTERIOR_BEGIN(h)
        double sum = 0.0, sumx = 0.0, sumy = 0.0, sumz = 0.0;
        for (int n = 0; n < q; n++) {
            int64_t ndxn = ndx * q + n;
            sum += h->f[ndxn];
            sumx += h->f[ndxn] * h->ex[n];
            sumy += h->f[ndxn] * h->ey[n];
            if (h->dim == 3) sumz += h->f[ndxn] * h->ez[n];
        }
        h->r[ndx] = sum;
        h->u[ndx] = sumx / h->r[ndx];
        h->v[ndx] = sumy / h->r[ndx];
        if (h->dim == 3) h->w[ndx] = sumz / h->r[ndx];
    FOR_INTERIOR_END
}

/* Collision into fs, then periodic ghosts of fs, then pull streaming:
 * LBfluid3D::collideStreamUpdate, LBfluid3D.cpp:252-304;
 * LBfluid2D::updateFluid_SC, LBfluid2D.cpp:235-272. */
static void collide_stream(scmp_oracle* h)
{
    const int q = h->q;
    const double tau = h->tau;
    FOR_INTERIOR_BEGIN(h)
        double ueq, veq, weq, usq;
        shifted_velocity(h, ndx, &ueq, &veq, &weq, &usq);
        for (int n = 0; n < q; n++) {
            int64_t ndxn = ndx * q + n;
            double feq = feq_n(h, n, h->r[ndx], ueq, veq, weq, usq);
            h->fs[ndxn] = h->f[ndxn] - (h->f[ndxn] - feq) / tau;
        }
    FOR_INTERIOR_END
    ghost_fill(h, h->fs, q);
    FOR_INTERIOR_BEGIN(h)
        for (int n = 0; n < q; n++) {
            int64_t nbr = nd(h, i - h->exi[n], j - h->eyi[n],
                             h->dim == 3 ? k - h->ezi[n] : 0);
            h->f[ndx * q + n] = h->fs[nbr * q + n];
        }
    FOR_INTERIOR_END
}

/* Halfway bounce-back, applied after streaming from the same node's
 * post-collision values.  3-D: LBfluid3D::bounceBackWallsZdir,
 * LBfluid3D.cpp:316-347 (rows k=2 and k=nz-1).  2-D:
 * LBfluid2D::bounceBackWallsYdir, LBfluid2D.cpp:363-388 (rows j=2, ny-1). */
static void bounce_back(scmp_oracle* h)
{
    const int q = h->q;
    if (h->dim == 3) {
        static const int lo_dst[5] = {5, 9, 11, 16, 18};
        static const int lo_src[5] = {6, 10, 12, 15, 17};
        for (int64_t i = 1; i < h->nx + 1; i++) {
            for (int64_t j = 1; j < h->ny + 1; j++) {
                int64_t a = nd(h, i, j, 2), b = nd(h, i, j, h->nz - 1);
                for (int m = 0; m < 5; m++)
                    h->f[a * q + lo_dst[m]] = h->fs[a * q + lo_src[m]];
                for (int m = 0; m < 5; m++)
                    h->f[b * q + lo_src[m]] = h->fs[b * q + lo_dst[m]];
            }
        }
    } else {
        static const int lo_dst[3] = {3, 5, 7};
        static const int lo_src[3] = {4, 6, 8};
        for (int64_t i = 1; i < h->nx + 1; i++) {
            int64_t a = nd(h, i, 2, 0), b = nd(h, i, h->ny - 1, 0);
            for (int m = 0; m < 3; m++)
                h->f[a * q + lo_dst[m]] = h->fs[a * q + lo_src[m]];
            for (int m = 0; m < 3; m++)
                h->f[b * q + lo_src[m]] = h->fs[b * q + lo_dst[m]];
        }
    }
}

void scmp_oracle_set_rho(scmp_oracle* h, const double* rho0)
{
    int64_t c = 0;
    for (int i = 1; i < h->nx + 1; i++)
        for (int j = 1; j < h->ny + 1; j++)
            for (int k = (h->dim == 3 ? 1 : 0); k < (h->dim == 3 ? h->nz + 1 : 1); k++) {
                int64_t ndx = nd(h, i, j, k);
                h->r[ndx] = rho0[c++];
                h->u[ndx] = 0.0;
                h->v[ndx] = 0.0;
                h->w[ndx] = 0.0;
            }
}

void scmp_oracle_init(scmp_oracle* h)
{
    ghost_fill(h, h->r, 1);          /* scmp3D.cpp:78 */
    shan_chen_forces(h);             /* scmp3D.cpp:79 */
    set_f_to_feq(h);                 /* scmp3D.cpp:80 */
}

void scmp_oracle_step(scmp_oracle* h, int nsteps)
{
    for (int s = 0; s < nsteps; s++) {
        if (h->order == 3) {
            /* scmp3D.cpp:90-119; walls as mcmp3DFilm.cpp:132-140 */
            macros(h);
            ghost_fill(h, h->r, 1);
            shan_chen_forces(h);
            collide_stream(h);
            if (h->walls) bounce_back(h);
        } else {
            /* mcmp2D.cpp:111-121 -> LBfluid2D.cpp:227-312 */
            shan_chen_forces(h);
            collide_stream(h);
            if (h->walls) bounce_back(h);
            macros(h);
            ghost_fill(h, h->r, 1);
        }
    }
}

void scmp_oracle_get(const scmp_oracle* h, double* rho, double* u, double* v,
                     double* w, double* f)
{
    int64_t c = 0;
    for (int i = 1; i < h->nx + 1; i++)
        for (int j = 1; j < h->ny + 1; j++)
            for (int k = (h->dim == 3 ? 1 : 0); k < (h->dim == 3 ? h->nz + 1 : 1); k++) {
                int64_t ndx = nd(h, i, j, k);
                if (rho) rho[c] = h->r[ndx];
                if (u) u[c] = h->u[ndx];
                if (v) v[c] = h->v[ndx];
                if (w) w[c] = (h->dim == 3) ? h->w[ndx] : 0.0;
                if (f) memcpy(f + c * h->q, h->f + ndx * h->q,
                              (size_t)h->q * sizeof(double));
                c++;
            }
}

void scmp_oracle_get_force(const scmp_oracle* h, double* fx, double* fy,
                           double* fz)
{
    int64_t c = 0;
    for (int i = 1; i < h->nx + 1; i++)
        for (int j = 1; j < h->ny + 1; j++)
            for (int k = (h->dim == 3 ? 1 : 0); k < (h->dim == 3 ? h->nz + 1 : 1); k++) {
                int64_t ndx = nd(h, i, j, k);
                if (fx) fx[c] = h->fx[ndx];
                if (fy) fy[c] = h->fy[ndx];
                if (fz) fz[c] = (h->dim == 3) ? h->fz[ndx] : 0.0;
                c++;
            }
}

void scmp_oracle_moments(const scmp_oracle* h, double* rho, double* u,
                         double* v, double* w)
{
    const int q = h->q;
    int64_t c = 0;
    for (int i = 1; i < h->nx + 1; i++)
        for (int j = 1; j < h->ny + 1; j++)
            for (int k = (h->dim == 3 ? 1 : 0); k < (h->dim == 3 ? h->nz + 1 : 1); k++) {
                int64_t ndx = nd(h, i, j, k);
                double sum = 0.0, sumx = 0.0, sumy = 0.0, sumz = 0.0;
                for (int n = 0; n < q; n++) {
                    double fv = h->f[ndx * q + n];
                    sum += fv;
                    sumx += fv * h->ex[n];
                    sumy += fv * h->ey[n];
                    sumz += fv * h->ez[n];
                }
                if (rho) rho[c] = sum;
                if (u) u[c] = sumx / sum;
                if (v) v[c] = sumy / sum;
                if (w) w[c] = (h->dim == 3) ? sumz / sum : 0.0;
                c++;
            }
}

void scmp_oracle_set_f(scmp_oracle* h, const double* f)
{
    int64_t c = 0;
    for (int i = 1; i < h->nx + 1; i++)
        for (int j = 1; j < h->ny + 1; j++)
            for (int k = (h->dim == 3 ? 1 : 0); k < (h->dim == 3 ? h->nz + 1 : 1); k++) {
                int64_t ndx = nd(h, i, j, k);
                memcpy(h->f + ndx * h->q, f + c * h->q,
                       (size_t)h->q * sizeof(double));
                c++;
            }
    /* The r,u arrays become the moments of the new f, as after a step of
     * the LBfluid2D order (LBfluid2D.cpp:285-310). */
    macros(h);
    ghost_fill(h, h->r, 1);
}

/* Seeded initial densities (the reference seeds rand() from time(), which is
 * not reproducible; these generators define the seeded inputs instead). */

/* std::mt19937(seed): the Mersenne Twister of Matsumoto & Nishimura (1998),
 * the same generator as src/utils/rand.cpp:24-78. */
typedef struct { uint32_t mt[624]; int idx; } mt19937_state;

static void mt_seed(mt19937_state* s, uint32_t seed)
{
    s->mt[0] = seed;
    for (int i = 1; i < 624; i++)
        s->mt[i] = 1812433253u * (s->mt[i - 1] ^ (s->mt[i - 1] >> 30)) + (uint32_t)i;
    s->idx = 624;
}

static uint32_t mt_next(mt19937_state* s)
{
    if (s->idx >= 624) {
        for (int i = 0; i < 624; i++) {
            uint32_t y = (s->mt[i] & 0x80000000u) | (s->mt[(i + 1) % 624] & 0x7fffffffu);
            uint32_t v = s->mt[(i + 397) % 624] ^ (y >> 1);
            if (y & 1u) v ^= 0x9908b0dfu;
            s->mt[i] = v;
        }
        s->idx = 0;
    }
    uint32_t y = s->mt[s->idx++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

/* rho = mean + amp*(x/4294967295.0 - 0.5), x = successive mt19937 outputs in
 * reference loop order; the mapping of scmp3D.cpp:67-69 with rand()/RAND_MAX
 * replaced by Rand::uniform (rand.cpp:81-84). */
void scmp_oracle_noise_rho(uint32_t seed, int64_t n, double mean, double amp,
                           double* out)
{
    mt19937_state s;
    mt_seed(&s, seed);
    for (int64_t c = 0; c < n; c++) {
        double x = (double)mt_next(&s) / 4294967295.0;
        out[c] = mean + amp * (x - 0.5);
    }
}

/* Raw generator output, for the known-answer test of mt19937. */
void scmp_oracle_mt19937(uint32_t seed, int64_t n, uint32_t* out)
{
    mt19937_state s;
    mt_seed(&s, seed);
    for (int64_t c = 0; c < n; c++) out[c] = mt_next(&s);
}

/* Droplet: rho = (rl+rv)/2 - (rl-rv)/2 * tanh(2(r-R)/W), r measured from the
 * geometric centre ((N-1)/2 per axis) without periodic images.  dims has 2
 * or 3 entries; nz = 1 for 2-D. */
void scmp_oracle_droplet_rho(int nx, int ny, int nz, double R, double W,
                             double rl, double rv, double* out)
{
    double cx = 0.5 * (nx - 1), cy = 0.5 * (ny - 1), cz = 0.5 * (nz - 1);
    int64_t c = 0;
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) {
                double dx = i - cx, dy = j - cy, dz = k - cz;
                double r = sqrt(dx * dx + dy * dy + dz * dz);
                out[c++] = 0.5 * (rl + rv) - 0.5 * (rl - rv) * tanh(2.0 * (r - R) / W);
            }
}

/* ---- two components: mcmp2D / mcmp3D ---------------------------------------
 * Two single-component states (their own tau, r/u/v/w, force arrays, f) and
 * one coupling G; the Shan-Chen force acts between the components only and
 * both collide with the common velocity.  mcmp2D.cpp:10-212,
 * mcmp3D.cpp:10-233. */
struct mcmp_oracle {
    scmp_oracle* c[2];
    double G;
    int order;
    int film;              /* mcmp2DFilm / mcmp3DFilm: walls on the last axis */
    double rho_sol[2];     /* wall densities of A and B */
};

mcmp_oracle* mcmp_oracle_create(int q, int nx, int ny, int nz, double tauA, double tauB, double G,
                                int order, int nthreads)
{
    mcmp_oracle* m = (mcmp_oracle*)calloc(1, sizeof(*m));
    if (!m) return NULL;
    m->c[0] = scmp_oracle_create(q, nx, ny, nz, tauA, 0.0, order, 0, nthreads);
    m->c[1] = scmp_oracle_create(q, nx, ny, nz, tauB, 0.0, order, 0, nthreads);
    m->G = G;
    m->order = order;
    if (!m->c[0] || !m->c[1]) {
        mcmp_oracle_destroy(m);
        return NULL;
    }
    return m;
}

void mcmp_oracle_destroy(mcmp_oracle* m)
{
    if (!m) return;
    scmp_oracle_destroy(m->c[0]);
    scmp_oracle_destroy(m->c[1]);
    free(m);
}

/* mcmp2D::calculateShanChenForces (mcmp2D.cpp:145-201) and its 3-D form
 * (mcmp3D.cpp:155-225): psi of both components at the 19 (9) neighbours in
 * stencil order, the cross forces -G psi_A S_B and -G psi_B S_A, and the
 * common velocity u' = (u_A r_A/tau_A + u_B r_B/tau_B) / (r_A/tau_A +
 * r_B/tau_B) written into both components' u arrays. */
static void mcmp_forces(mcmp_oracle* m)
{
    scmp_oracle* a = m->c[0];
    scmp_oracle* b = m->c[1];
    const scmp_oracle* h = a;
    const int q = a->q;
    const double G = m->G;
    FOR_INTERIOR_BEGIN(h)
        double rhoA = a->r[ndx];
        double rhoB = b->r[ndx];
        double psi_rhoA = psi(a, rhoA);
        double psi_rhoB = psi(b, rhoB);
        double GsumxA = 0.0, GsumyA = 0.0, GsumzA = 0.0;
        double GsumxB = 0.0, GsumyB = 0.0, GsumzB = 0.0;
        for (int n = 0; n < q; n++) {
            int64_t nbr = nd(h, i + h->exi[n], j + h->eyi[n], h->dim == 3 ? k + h->ezi[n] : 0);
            double psi_rhoA_nbr = psi(a, a->r[nbr]);
            double psi_rhoB_nbr = psi(b, b->r[nbr]);
            GsumxA += h->wa[n] * h->ex[n] * psi_rhoA_nbr;
            GsumyA += h->wa[n] * h->ey[n] * psi_rhoA_nbr;
            GsumzA += h->wa[n] * h->ez[n] * psi_rhoA_nbr;
            GsumxB += h->wa[n] * h->ex[n] * psi_rhoB_nbr;
            GsumyB += h->wa[n] * h->ey[n] * psi_rhoB_nbr;
            GsumzB += h->wa[n] * h->ez[n] * psi_rhoB_nbr;
        }
        a->fx[ndx] = -G * psi_rhoA * GsumxB;
        a->fy[ndx] = -G * psi_rhoA * GsumyB;
        b->fx[ndx] = -G * psi_rhoB * GsumxA;
        b->fy[ndx] = -G * psi_rhoB * GsumyA;
        if (h->dim == 3) {
            a->fz[ndx] = -G * psi_rhoA * GsumzB;
            b->fz[ndx] = -G * psi_rhoB * GsumzA;
        }
        /* LBfluid2D::getURhoDivTau etc. (LBfluid2D.cpp:179-192): u*r/tau, r/tau */
        double rdivtauA = a->r[ndx] / a->tau;
        double rdivtauB = b->r[ndx] / b->tau;
        double uP = (a->u[ndx] * a->r[ndx] / a->tau + b->u[ndx] * b->r[ndx] / b->tau) /
                    (rdivtauA + rdivtauB);
        double vP = (a->v[ndx] * a->r[ndx] / a->tau + b->v[ndx] * b->r[ndx] / b->tau) /
                    (rdivtauA + rdivtauB);
        a->u[ndx] = uP; a->v[ndx] = vP;
        b->u[ndx] = uP; b->v[ndx] = vP;
        if (h->dim == 3) {
            double wP = (a->w[ndx] * a->r[ndx] / a->tau + b->w[ndx] * b->r[ndx] / b->tau) /
                        (rdivtauA + rdivtauB);
            a->w[ndx] = wP;
            b->w[ndx] = wP;
        }
    FOR_INTERIOR_END
}

/* mcmp2DFilm::calculateShanChenForces (mcmp2DFilm.cpp:149-219) and
 * mcmp3DFilm's (mcmp3DFilm.cpp:167-253): rows 1 and n of the last axis are
 * walls; the loop skips them (their force and u arrays are left as they
 * are), and a neighbour in a wall row contributes psi(rho_sol) of the
 * component instead of psi of its density. */
static void mcmp_forces_film(mcmp_oracle* m)
{
    scmp_oracle* a = m->c[0];
    scmp_oracle* b = m->c[1];
    const scmp_oracle* h = a;
    const int q = a->q;
    const double G = m->G;
    const int nl = (h->dim == 3) ? h->nz : h->ny;
    _Pragma("omp parallel for num_threads(h->nthreads) schedule(static)")
    for (int i = 1; i < h->nx + 1; i++) {
        for (int j = (h->dim == 3 ? 1 : 2); j < (h->dim == 3 ? h->ny + 1 : h->ny); j++) {
            for (int k = (h->dim == 3 ? 2 : 0); k < (h->dim == 3 ? h->nz : 1); k++) {
                int64_t ndx = nd(h, i, j, k);
                double rhoA = a->r[ndx];
                double rhoB = b->r[ndx];
                double psi_rhoA = psi(a, rhoA);
                double psi_rhoB = psi(b, rhoB);
                double GsumxA = 0.0, GsumyA = 0.0, GsumzA = 0.0;
                double GsumxB = 0.0, GsumyB = 0.0, GsumzB = 0.0;
                for (int n = 0; n < q; n++) {
                    int inbr = i + h->exi[n], jnbr = j + h->eyi[n];
                    int knbr = (h->dim == 3) ? k + h->ezi[n] : 0;
                    int lnbr = (h->dim == 3) ? knbr : jnbr;
                    double psi_rhoA_nbr, psi_rhoB_nbr;
                    if (lnbr == 1 || lnbr == nl) {
                        psi_rhoA_nbr = psi(a, m->rho_sol[0]);
                        psi_rhoB_nbr = psi(b, m->rho_sol[1]);
                    } else {
                        int64_t nbr = nd(h, inbr, jnbr, knbr);
                        psi_rhoA_nbr = psi(a, a->r[nbr]);
                        psi_rhoB_nbr = psi(b, b->r[nbr]);
                    }
                    GsumxA += h->wa[n] * h->ex[n] * psi_rhoA_nbr;
                    GsumyA += h->wa[n] * h->ey[n] * psi_rhoA_nbr;
                    GsumzA += h->wa[n] * h->ez[n] * psi_rhoA_nbr;
                    GsumxB += h->wa[n] * h->ex[n] * psi_rhoB_nbr;
                    GsumyB += h->wa[n] * h->ey[n] * psi_rhoB_nbr;
                    GsumzB += h->wa[n] * h->ez[n] * psi_rhoB_nbr;
                }
                a->fx[ndx] = -G * psi_rhoA * GsumxB;
                a->fy[ndx] = -G * psi_rhoA * GsumyB;
                b->fx[ndx] = -G * psi_rhoB * GsumxA;
                b->fy[ndx] = -G * psi_rhoB * GsumyA;
                if (h->dim == 3) {
                    a->fz[ndx] = -G * psi_rhoA * GsumzB;
                    b->fz[ndx] = -G * psi_rhoB * GsumzA;
                }
                double rdivtauA = a->r[ndx] / a->tau;
                double rdivtauB = b->r[ndx] / b->tau;
                double uP = (a->u[ndx] * a->r[ndx] / a->tau + b->u[ndx] * b->r[ndx] / b->tau) /
                            (rdivtauA + rdivtauB);
                double vP = (a->v[ndx] * a->r[ndx] / a->tau + b->v[ndx] * b->r[ndx] / b->tau) /
                            (rdivtauA + rdivtauB);
                a->u[ndx] = uP; a->v[ndx] = vP;
                b->u[ndx] = uP; b->v[ndx] = vP;
                if (h->dim == 3) {
                    double wP = (a->w[ndx] * a->r[ndx] / a->tau + b->w[ndx] * b->r[ndx] / b->tau) /
                                (rdivtauA + rdivtauB);
                    a->w[ndx] = wP;
                    b->w[ndx] = wP;
                }
            }
        }
    }
}

static void forces_of(mcmp_oracle* m)
{
    if (m->film) mcmp_forces_film(m);
    else mcmp_forces(m);
}

void mcmp_oracle_set_film(mcmp_oracle* m, double rhoA_sol, double rhoB_sol)
{
    m->film = 1;
    m->rho_sol[0] = rhoA_sol;
    m->rho_sol[1] = rhoB_sol;
    m->c[0]->walls = 1;
    m->c[1]->walls = 1;
}

void mcmp_oracle_set_rho(mcmp_oracle* m, const double* rhoA, const double* rhoB)
{
    /* mcmp2D.cpp:67-81, mcmp3D.cpp:67-85 (velocities zero) */
    scmp_oracle_set_rho(m->c[0], rhoA);
    scmp_oracle_set_rho(m->c[1], rhoB);
}

void mcmp_oracle_init(mcmp_oracle* m)
{
    /* mcmp2D.cpp:83-87 / mcmp3D.cpp:87-91 */
    ghost_fill(m->c[0], m->c[0]->r, 1);
    ghost_fill(m->c[1], m->c[1]->r, 1);
    forces_of(m);
    set_f_to_feq(m->c[0]);
    set_f_to_feq(m->c[1]);
}

void mcmp_oracle_step(mcmp_oracle* m, int nsteps)
{
    for (int s = 0; s < nsteps; s++) {
        if (m->order == 3) {
            /* mcmp3D::updateLatticeBoltzmann, mcmp3D.cpp:97-123; the film
             * adds bounce-back on both after streaming, mcmp3DFilm.cpp:132-140 */
            for (int c = 0; c < 2; c++) {
                macros(m->c[c]);
                ghost_fill(m->c[c], m->c[c]->r, 1);
            }
            forces_of(m);
            collide_stream(m->c[0]);
            collide_stream(m->c[1]);
            if (m->film) {
                bounce_back(m->c[0]);
                bounce_back(m->c[1]);
            }
        } else {
            /* mcmp2D::updateLatticeBoltzmann, mcmp2D.cpp:97-123, with
             * LBfluid2D::updateFluid_SC per component (LBfluid2D.cpp:227-312);
             * the film sets yBC = 1 (mcmp2DFilm.cpp:119-123) */
            forces_of(m);
            for (int c = 0; c < 2; c++) {
                collide_stream(m->c[c]);
                if (m->film) bounce_back(m->c[c]);
                macros(m->c[c]);
                ghost_fill(m->c[c], m->c[c]->r, 1);
            }
        }
    }
}

void mcmp_oracle_get(const mcmp_oracle* m, int comp, double* rho, double* u, double* v, double* w,
                     double* f)
{
    scmp_oracle_get(m->c[comp & 1], rho, u, v, w, f);
}

void mcmp_oracle_get_force(const mcmp_oracle* m, int comp, double* fx, double* fy, double* fz)
{
    scmp_oracle_get_force(m->c[comp & 1], fx, fy, fz);
}

void mcmp_oracle_moments(const mcmp_oracle* m, int comp, double* rho, double* u, double* v,
                         double* w)
{
    scmp_oracle_moments(m->c[comp & 1], rho, u, v, w);
}

/* ---- immersed-boundary capsule (capsule2D) ----------------------------------
 * capsule2D.cpp:84-150 with IBcapsules2D.cpp:130-484: one D2Q9 fluid in the
 * LBfluid2D step order with bounce-back walls in y (yBC = 1), no Shan-Chen
 * force; its force arrays hold a body force plus the membrane forces of the
 * IB nodes spread with the bilinear kernel; the nodes move with the
 * interpolated Shan-Chen velocity u* = u + F/(2 rho). */
struct capsule_oracle {
    scmp_oracle* fl;
    int nnodes;
    double kstiff, dt, body_fx, body_fy;
    double *x, *y, *x0, *y0, *vx, *vy, *nfx, *nfy;
};

capsule_oracle* capsule_oracle_create(int nx, int ny, double tau, int nnodes, double kstiff,
                                      double dt, double body_fx, double body_fy)
{
    if (nnodes < 1) return NULL;
    capsule_oracle* c = (capsule_oracle*)calloc(1, sizeof(*c));
    if (!c) return NULL;
    c->fl = scmp_oracle_create(9, nx, ny, 1, tau, 0.0, 2, 1, 1);
    c->nnodes = nnodes;
    c->kstiff = kstiff;
    c->dt = dt;
    c->body_fx = body_fx;
    c->body_fy = body_fy;
    double** arrs[8] = {&c->x, &c->y, &c->x0, &c->y0, &c->vx, &c->vy, &c->nfx, &c->nfy};
    int ok = c->fl != NULL;
    for (int a = 0; a < 8; a++) {
        *arrs[a] = (double*)calloc((size_t)nnodes, sizeof(double));
        ok = ok && *arrs[a];
    }
    if (!ok) {
        capsule_oracle_destroy(c);
        return NULL;
    }
    return c;
}

void capsule_oracle_destroy(capsule_oracle* c)
{
    if (!c) return;
    scmp_oracle_destroy(c->fl);
    free(c->x); free(c->y); free(c->x0); free(c->y0);
    free(c->vx); free(c->vy); free(c->nfx); free(c->nfy);
    free(c);
}

void capsule_oracle_set_rho(capsule_oracle* c, const double* rho0)
{
    /* capsule2D.cpp:92-107: rho, u = v = 0 and fx = fy = 0 on the interior */
    scmp_oracle_set_rho(c->fl, rho0);
}

void capsule_oracle_set_nodes(capsule_oracle* c, const double* x, const double* y)
{
    for (int n = 0; n < c->nnodes; n++) {
        c->x[n] = x[n];
        c->y[n] = y[n];
        c->x0[n] = x[n];
        c->y0[n] = y[n];
    }
}

void capsule_oracle_move_nodes(capsule_oracle* c, const double* x, const double* y)
{
    for (int n = 0; n < c->nnodes; n++) {
        c->x[n] = x[n];
        c->y[n] = y[n];
    }
}

void capsule_oracle_circle(capsule_oracle* c, double xo, double yo, double rad)
{
    /* IBcapsules2D::initCircularCapsule, IBcapsules2D.cpp:130-141 */
    double theta = 0.0;
    const double dTheta = (2.0 * 3.14159265) / (double)c->nnodes;
    for (int n = 0; n < c->nnodes; n++) {
        c->x[n] = xo + rad * cos(theta);
        c->y[n] = yo + rad * sin(theta);
        c->x0[n] = c->x[n];
        c->y0[n] = c->y[n];
        theta += dTheta;
    }
}

void capsule_oracle_init(capsule_oracle* c)
{
    /* capsule2D.cpp:118: setFtoFeq with the zero forces of the start */
    set_f_to_feq(c->fl);
}

/* Footprint origin of an IB node: i0 = int(floor(x)) + 1 and the same in y,
 * ghost-shell indices (IBcapsules2D.cpp:238-239, 352-353).  Returns 0 when
 * the reference skips the node, isMyNode(i0) false on one rank
 * (IBcapsules2D.cpp:418-426: 0 <= i0 <= nx), which covers every non-finite
 * x.  A node whose y leaves [-1, ny) would index outside the ghost shell in
 * the reference (no y check there); such a node is skipped here. */
static int ib_origin(const scmp_oracle* h, double x, double y, int* i0, int* j0)
{
    if (!(x >= -1.0 && x < (double)h->nx)) return 0;
    if (!(y >= -1.0 && y < (double)h->ny)) return 0;
    *i0 = (int)floor(x) + 1;
    *j0 = (int)floor(y) + 1;
    return 1;
}

static void capsule_step(capsule_oracle* c)
{
    scmp_oracle* h = c->fl;
    /* IBcapsules2D::computeNodeForces, IBcapsules2D.cpp:477-484 */
    for (int n = 0; n < c->nnodes; n++) {
        c->nfx[n] = -c->kstiff * (c->x[n] - c->x0[n]);
        c->nfy[n] = -c->kstiff * (c->y[n] - c->y0[n]);
    }
    /* LBfluid2D::zeroForces (interior only, LBfluid2D.cpp:137-147), then the
     * body force on rows j = 2 .. ny-1 (capsule2D.cpp:128-136) */
    for (int i = 1; i < h->nx + 1; i++)
        for (int j = 1; j < h->ny + 1; j++) {
            int64_t ndx = nd(h, i, j, 0);
            h->fx[ndx] = 0.0;
            h->fy[ndx] = 0.0;
        }
    for (int i = 1; i < h->nx + 1; i++)
        for (int j = 2; j < h->ny; j++) {
            int64_t ndx = nd(h, i, j, 0);
            h->fx[ndx] = c->body_fx;
            h->fy[ndx] = c->body_fy;
        }
    /* IBcapsules2D::extrapolateForce, IBcapsules2D.cpp:339-386: bilinear
     * spread onto i0..i0+1 (inside 1..nx) and j0..j0+1 (ghost rows
     * included: they are never zeroed, so what lands there adds up over the
     * steps) */
    for (int n = 0; n < c->nnodes; n++) {
        int i0, j0;
        if (!ib_origin(h, c->x[n], c->y[n], &i0, &j0)) continue;
        for (int i = i0; i <= i0 + 1; i++) {
            if (i < 1 || i > h->nx) continue;
            for (int j = j0; j <= j0 + 1; j++) {
                int64_t ndx = nd(h, i, j, 0);
                double rx = c->x[n] - (double)(i - 1);
                double ry = c->y[n] - (double)(j - 1);
                double del = (1.0 - fabs(rx)) * (1.0 - fabs(ry));
                h->fx[ndx] += del * c->nfx[n];
                h->fy[ndx] += del * c->nfy[n];
            }
        }
    }
    /* LBfluid2D::updateFluid_SC with xBC = 0, yBC = 1 and no ghost rho
     * (capsule2D.cpp:142-145, LBfluid2D.cpp:227-312) */
    collide_stream(h);
    bounce_back(h);
    macros(h);
    /* IBcapsules2D::interpolateVelocity, IBcapsules2D.cpp:218-287 (the
     * Allreduce of one rank is the identity), with getUStar
     * (LBfluid2D.cpp:164-172); ghost rows hold r = u = v = 0 */
    for (int n = 0; n < c->nnodes; n++) {
        c->vx[n] = 0.0;
        c->vy[n] = 0.0;
        int i0, j0;
        if (!ib_origin(h, c->x[n], c->y[n], &i0, &j0)) continue;
        for (int i = i0; i <= i0 + 1; i++) {
            if (i < 1 || i > h->nx) continue;
            for (int j = j0; j <= j0 + 1; j++) {
                int64_t ndx = nd(h, i, j, 0);
                double rx = c->x[n] - (double)(i - 1);
                double ry = c->y[n] - (double)(j - 1);
                double del = (1.0 - fabs(rx)) * (1.0 - fabs(ry));
                double us = h->u[ndx] + 0.5 * h->fx[ndx] / h->r[ndx];
                double vs = h->v[ndx] + 0.5 * h->fy[ndx] / h->r[ndx];
                c->vx[n] += del * us;
                c->vy[n] += del * vs;
            }
        }
    }
    /* IBcapsules2D::updateNodePositions, IBcapsules2D.cpp:463-469 */
    for (int n = 0; n < c->nnodes; n++) {
        c->x[n] += c->vx[n] * c->dt;
        c->y[n] += c->vy[n] * c->dt;
    }
}

void capsule_oracle_step(capsule_oracle* c, int nsteps)
{
    for (int s = 0; s < nsteps; s++) capsule_step(c);
}

void capsule_oracle_get(const capsule_oracle* c, double* rho, double* u, double* v, double* f,
                        double* fx, double* fy)
{
    scmp_oracle_get(c->fl, rho, u, v, NULL, f);
    scmp_oracle_get_force(c->fl, fx, fy, NULL);
}

void capsule_oracle_get_nodes(const capsule_oracle* c, double* x, double* y, double* vx, double* vy,
                              double* fx, double* fy)
{
    for (int n = 0; n < c->nnodes; n++) {
        if (x) x[n] = c->x[n];
        if (y) y[n] = c->y[n];
        if (vx) vx[n] = c->vx[n];
        if (vy) vy[n] = c->vy[n];
        if (fx) fx[n] = c->nfx[n];
        if (fy) fy[n] = c->nfy[n];
    }
}

void capsule_oracle_get_ghost_force(const capsule_oracle* c, double* gfx, double* gfy)
{
    /* rows j = 0 and j = ny+1 of the force arrays, i = 1..nx */
    const scmp_oracle* h = c->fl;
    for (int s = 0; s < 2; s++)
        for (int i = 1; i < h->nx + 1; i++) {
            int64_t ndx = nd(h, i, s == 0 ? 0 : h->ny + 1, 0);
            gfx[s * h->nx + (i - 1)] = h->fx[ndx];
            gfy[s * h->nx + (i - 1)] = h->fy[ndx];
        }
}
