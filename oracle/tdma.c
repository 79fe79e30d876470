/*
 * oracle/tdma.c -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement, in plain C and FP64, of EAF3D's line-by-line TDMA
 * smoother.  It is the checker for the HIP line solver: only tests and the
 * benchmark's CPU-baseline leg load it.
 *
 *   oracle_tdma     the Thomas algorithm `tdma` (src/mod_solver_3d.f90:22-46):
 *                   forward elimination with the pivot clamped to
 *                   sign(SMALL, m) when |m| < SMALL, then back substitution
 *   oracle_tdma_3d  `tdma_3d_mpi` on a serial mesh (src/mod_solver_3d.f90:
 *                   252-366): per sweep, the theta halo copy, the r lines
 *                   (k outer, j inner), the halo copy, the theta lines
 *                   (k outer, i inner), the halo copy, the z lines (j outer,
 *                   i inner).  Each line's right-hand side takes the other
 *                   directions' neighbours from phi as it stands (updated
 *                   lines before, old lines after, halos as last copied) and
 *                   its two end neighbours from the halos; no halo copy
 *                   follows the last z pass.
 *
 * Arrays use the reference's padded Fortran layout, bounds (-1:n+2), r
 * fastest.  Expressions keep the reference's operand order, built with
 * -ffp-contract=off.  Fortran's sign(SMALL, m) is copysign(SMALL, m).
 *
 * Parity is pinned by tests/test_tdma_oracle.py: the reference's own
 * known-answer test (tests/unit/test_tdma.f90: x = sin(i), b = 4, a = c = 1,
 * n = 50, error below 1e-12), scipy's banded solver, the pivot clamp, each
 * pass against an independent sparse block Gauss-Seidel statement of the
 * same line order, and convergence of many sweeps to scipy's direct solve.
 */
#include <math.h>
#include <stdlib.h>

#define SMALL 1.0e-30                 /* src/mod_constants.f90:53 */

/* x(1:n) solving a(i) x(i-1) + b(i) x(i) + c(i) x(i+1) = d(i); cp and dp
 * are caller scratch of n doubles. */
static void tdma_ws(const double* a, const double* b, const double* c, const double* d, double* x, int n,
                    double* cp, double* dp) {
    double b1 = b[0];
    if (fabs(b1) < SMALL) b1 = copysign(SMALL, b1);
    cp[0] = c[0] / b1;
    dp[0] = d[0] / b1;
    for (int i = 1; i < n; ++i) {
        double m = b[i] - a[i] * cp[i - 1];
        if (fabs(m) < SMALL) m = copysign(SMALL, m);
        cp[i] = c[i] / m;
        dp[i] = (d[i] - a[i] * dp[i - 1]) / m;
    }
    x[n - 1] = dp[n - 1];
    for (int i = n - 2; i >= 0; --i) x[i] = dp[i] - cp[i] * x[i + 1];
}

int oracle_tdma(const double* a, const double* b, const double* c, const double* d, double* x, int n) {
    if (n < 1) return -1;
    double* w = malloc(sizeof(double) * 2 * (size_t)n);
    if (!w) return -1;
    tdma_ws(a, b, c, d, x, n, w, w + n);
    free(w);
    return 0;
}

typedef struct {
    int nr, nth, nz;
    long sy, sz;
} grid_t;

static inline long O(const grid_t* g, int i, int j, int k) {
    return (long)(i + 1) + g->sy * (j + 1) + g->sz * (k + 1);
}

/* src/mod_mpi_topology.f90:303-310, the serial periodic copy. */
static void theta_copy(const grid_t* g, double* a) {
    const int dst[4] = {0, -1, g->nth + 1, g->nth + 2};
    const int src[4] = {g->nth, g->nth - 1, 1, 2};
    for (int q = 0; q < 4; ++q)
        for (int k = -1; k <= g->nz + 2; ++k)
            for (int i = -1; i <= g->nr + 2; ++i) a[O(g, i, dst[q], k)] = a[O(g, i, src[q], k)];
}

int oracle_tdma_3d(int nr, int nth, int nz, const double* aW, const double* aE, const double* aS, const double* aN,
                   const double* aB, const double* aT, const double* aP, const double* Su, double* phi, int n_sweep) {
    if (nr < 1 || nth < 1 || nz < 1) return -1;
    grid_t g;
    g.nr = nr; g.nth = nth; g.nz = nz;
    g.sy = nr + 4; g.sz = (long)(nr + 4) * (nth + 4);
    int nmax = nr > nth ? nr : nth;
    if (nz > nmax) nmax = nz;
    double* w = malloc(sizeof(double) * 7 * (size_t)nmax);
    if (!w) return -1;
    double *a = w, *b = w + nmax, *c = w + 2 * nmax, *d = w + 3 * nmax, *x = w + 4 * nmax;
    double *cp = w + 5 * nmax, *dp = w + 6 * nmax;
    for (int sweep = 1; sweep <= n_sweep; ++sweep) {
        theta_copy(&g, phi);
        /* r lines (:296-317) */
        for (int k = 1; k <= nz; ++k)
            for (int j = 1; j <= nth; ++j) {
                for (int i = 1; i <= nr; ++i) {
                    const long o = O(&g, i, j, k);
                    a[i - 1] = -aW[o];
                    b[i - 1] = aP[o];
                    c[i - 1] = -aE[o];
                    d[i - 1] = Su[o] + aS[o] * phi[O(&g, i, j - 1, k)] + aN[o] * phi[O(&g, i, j + 1, k)];
                    d[i - 1] = d[i - 1] + aB[o] * phi[O(&g, i, j, k - 1)] + aT[o] * phi[O(&g, i, j, k + 1)];
                }
                d[0] = d[0] + aW[O(&g, 1, j, k)] * phi[O(&g, 0, j, k)];
                d[nr - 1] = d[nr - 1] + aE[O(&g, nr, j, k)] * phi[O(&g, nr + 1, j, k)];
                tdma_ws(a, b, c, d, x, nr, cp, dp);
                for (int i = 1; i <= nr; ++i) phi[O(&g, i, j, k)] = x[i - 1];
            }
        theta_copy(&g, phi);
        /* theta lines (:322-342) */
        for (int k = 1; k <= nz; ++k)
            for (int i = 1; i <= nr; ++i) {
                for (int j = 1; j <= nth; ++j) {
                    const long o = O(&g, i, j, k);
                    a[j - 1] = -aS[o];
                    b[j - 1] = aP[o];
                    c[j - 1] = -aN[o];
                    d[j - 1] = Su[o];
                    d[j - 1] = d[j - 1] + aW[o] * phi[O(&g, i - 1, j, k)] + aE[o] * phi[O(&g, i + 1, j, k)];
                    d[j - 1] = d[j - 1] + aB[o] * phi[O(&g, i, j, k - 1)] + aT[o] * phi[O(&g, i, j, k + 1)];
                }
                d[0] = d[0] + aS[O(&g, i, 1, k)] * phi[O(&g, i, 0, k)];
                d[nth - 1] = d[nth - 1] + aN[O(&g, i, nth, k)] * phi[O(&g, i, nth + 1, k)];
                tdma_ws(a, b, c, d, x, nth, cp, dp);
                for (int j = 1; j <= nth; ++j) phi[O(&g, i, j, k)] = x[j - 1];
            }
        theta_copy(&g, phi);
        /* z lines (:347-366) */
        for (int j = 1; j <= nth; ++j)
            for (int i = 1; i <= nr; ++i) {
                for (int k = 1; k <= nz; ++k) {
                    const long o = O(&g, i, j, k);
                    a[k - 1] = -aB[o];
                    b[k - 1] = aP[o];
                    c[k - 1] = -aT[o];
                    d[k - 1] = Su[o] + aS[o] * phi[O(&g, i, j - 1, k)] + aN[o] * phi[O(&g, i, j + 1, k)];
                    d[k - 1] = d[k - 1] + aW[o] * phi[O(&g, i - 1, j, k)] + aE[o] * phi[O(&g, i + 1, j, k)];
                }
                d[0] = d[0] + aB[O(&g, i, j, 1)] * phi[O(&g, i, j, 0)];
                d[nz - 1] = d[nz - 1] + aT[O(&g, i, j, nz)] * phi[O(&g, i, j, nz + 1)];
                tdma_ws(a, b, c, d, x, nz, cp, dp);
                for (int k = 1; k <= nz; ++k) phi[O(&g, i, j, k)] = x[k - 1];
            }
    }
    free(w);
    return 0;
}
