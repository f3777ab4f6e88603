/* sanitize_check.c -- drives the oracle (scmp_oracle.c) through every code
 * path the tests use, for the AddressSanitizer / UndefinedBehaviorSanitizer
 * build (make -C oracle sanitize).  TEST INFRASTRUCTURE ONLY.
 *
 * Cases: D3Q19 and D2Q9, both step orders, walls, tau != 1, one and three
 * x-slab threads, set_f restarts and the seeded input generators.  Besides
 * running clean under the sanitizers it checks two invariants: results do
 * not depend on the thread count (bit for bit), and periodic lattices
 * conserve mass.  Exit status 0 when every check holds. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "scmp_oracle.h"

static int failures = 0;

static void check(int ok, const char* what)
{
    if (!ok) {
        fprintf(stderr, "FAIL: %s\n", what);
        failures++;
    }
}

static double sum(const double* a, int64_t n)
{
    double s = 0.0;
    for (int64_t i = 0; i < n; i++) s += a[i];
    return s;
}

/* runs one case with 1 and 3 threads; compares rho and f bit for bit */
static void run_case(int q, int nx, int ny, int nz, double tau, double G, int order, int walls,
                     int steps, int droplet)
{
    const int64_t n = (int64_t)nx * ny * nz;
    double* rho0 = malloc(sizeof(double) * n);
    if (droplet) scmp_oracle_droplet_rho(nx, ny, nz, 0.25 * nx, 2.0, 2.0, 0.15, rho0);
    else scmp_oracle_noise_rho(7u, n, 0.7, 0.02, rho0);
    double *rho[2], *f[2];
    for (int t = 0; t < 2; t++) {
        scmp_oracle* h = scmp_oracle_create(q, nx, ny, nz, tau, G, order, walls, t == 0 ? 1 : 3);
        check(h != NULL, "create");
        if (!h) return;
        check(scmp_oracle_nodes(h) == n, "node count");
        scmp_oracle_set_rho(h, rho0);
        scmp_oracle_init(h);
        scmp_oracle_step(h, steps / 2);
        /* restart from a snapshot of the populations */
        double* snap = malloc(sizeof(double) * n * q);
        scmp_oracle_get(h, NULL, NULL, NULL, NULL, snap);
        scmp_oracle_set_f(h, snap);
        free(snap);
        scmp_oracle_step(h, steps - steps / 2);
        rho[t] = malloc(sizeof(double) * n);
        f[t] = malloc(sizeof(double) * n * q);
        double* u = malloc(sizeof(double) * n);
        double* v = malloc(sizeof(double) * n);
        double* w = malloc(sizeof(double) * n);
        scmp_oracle_get(h, rho[t], u, v, w, f[t]);
        scmp_oracle_get_force(h, u, v, w);
        scmp_oracle_moments(h, rho[t], u, v, w);
        free(u);
        free(v);
        free(w);
        scmp_oracle_destroy(h);
    }
    char what[160];
    snprintf(what, sizeof(what), "thread invariance q=%d %dx%dx%d order %d walls %d", q, nx, ny, nz,
             order, walls);
    check(memcmp(rho[0], rho[1], sizeof(double) * n) == 0 &&
              memcmp(f[0], f[1], sizeof(double) * n * q) == 0,
          what);
    if (!walls) {
        const double m0 = sum(rho0, n), m1 = sum(rho[0], n);
        snprintf(what, sizeof(what), "mass q=%d %dx%dx%d: %.17g vs %.17g", q, nx, ny, nz, m0, m1);
        check(fabs(m1 / m0 - 1.0) < 1e-12, what);
    }
    for (int t = 0; t < 2; t++) {
        free(rho[t]);
        free(f[t]);
    }
    free(rho0);
}

int main(void)
{
    uint32_t mt[4];
    scmp_oracle_mt19937(5489u, 4, mt);
    check(mt[0] == 3499211612u, "mt19937 known answer");
    run_case(19, 7, 6, 5, 1.0, -5.5, 3, 0, 6, 0);
    run_case(19, 7, 6, 5, 0.7, -5.5, 2, 1, 6, 0);
    run_case(19, 8, 8, 10, 1.3, -5.0, 3, 1, 5, 1);
    run_case(9, 13, 11, 1, 1.0, -5.0, 2, 0, 9, 0);
    run_case(9, 12, 10, 1, 0.8, -5.0, 3, 1, 9, 1);
    run_case(9, 3, 3, 1, 1.0, -4.5, 2, 0, 4, 0);
    if (failures) {
        fprintf(stderr, "%d check(s) failed\n", failures);
        return 1;
    }
    printf("oracle sanitize check: ok\n");
    return 0;
}
