/*
 * oracle/cg_omp.c -- TEST INFRASTRUCTURE ONLY (the multi-core CPU baseline).
 *
 * The same FP64 Jacobi-PCG as oracle_cg in cg_oracle.c (the restatement of
 * src/mod_solver_3d.f90:485-620), with its loops shared over OpenMP threads.
 * The passes are the reference's:
 *   - matvec with the theta halo copy (:588-605);
 *   - the dots (:607-618);
 *   - the phi/r/z/pv sweeps.
 * Only the dot products change: each is a parallel reduction, so the
 * summation order (and the last bits of the iterates) differs from the
 * sequential Fortran sum().
 *
 * The benchmark times it on the host cores beside the GPU.
 * tests/test_oracle.py checks it against the sequential oracle: the same
 * iteration count within one, and the same solution.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define SMALL 1.0e-30   /* src/mod_constants.f90:53 */

typedef struct {
    int nr, nth, nz;
    long sy, sz, n;
} grid_t;

static inline long off(const grid_t* g, int i, int j, int k) {
    return (long)(i + 1) + g->sy * (j + 1) + g->sz * (k + 1);
}

static void theta_halo(const grid_t* g, double* x) {
    const int dst[4] = {0, -1, g->nth + 1, g->nth + 2};
    const int src[4] = {g->nth, g->nth - 1, 1, 2};
    for (int p = 0; p < 4; ++p) {
#pragma omp parallel for schedule(static)
        for (int k = -1; k <= g->nz + 2; ++k)
            for (int i = -1; i <= g->nr + 2; ++i) x[off(g, i, dst[p], k)] = x[off(g, i, src[p], k)];
    }
}

static void matvec(const grid_t* g, const double* const* c, double* x, double* y) {
    const double *aW = c[0], *aE = c[1], *aS = c[2], *aN = c[3], *aB = c[4], *aT = c[5], *aP = c[6];
    theta_halo(g, x);
#pragma omp parallel for schedule(static)
    for (long o = 0; o < g->n; ++o) y[o] = 0.0;
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= g->nz; ++k)
        for (int j = 1; j <= g->nth; ++j)
            for (int i = 1; i <= g->nr; ++i) {
                const long o = off(g, i, j, k);
                y[o] = aP[o] * x[o] - aW[o] * x[o - 1] - aE[o] * x[o + 1] - aS[o] * x[o - g->sy] -
                       aN[o] * x[o + g->sy] - aB[o] * x[o - g->sz] - aT[o] * x[o + g->sz];
            }
}

static double dot(const grid_t* g, const double* a, const double* b) {
    double s = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : s)
    for (int k = 1; k <= g->nz; ++k)
        for (int j = 1; j <= g->nth; ++j)
            for (int i = 1; i <= g->nr; ++i) {
                const long o = off(g, i, j, k);
                s += a[o] * b[o];
            }
    return s;
}

/* Interior sweep: out = a + s*b  (out may alias a). */
static void axpy(const grid_t* g, double* out, const double* a, double s, const double* b) {
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= g->nz; ++k)
        for (int j = 1; j <= g->nth; ++j)
            for (int i = 1; i <= g->nr; ++i) {
                const long o = off(g, i, j, k);
                out[o] = a[o] + s * b[o];
            }
}

static void jacobi(const grid_t* g, double* z, const double* r, const double* aP) {
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= g->nz; ++k)
        for (int j = 1; j <= g->nth; ++j)
            for (int i = 1; i <= g->nr; ++i) {
                const long o = off(g, i, j, k);
                z[o] = r[o] / fmax(fabs(aP[o]), SMALL);
            }
}

/* Returns 0 on success, -1 if the work arrays cannot be allocated. */
int oracle_cg_omp(int nr, int nth, int nz,
                  const double* aW, const double* aE, const double* aS, const double* aN,
                  const double* aB, const double* aT, const double* aP, const double* Su,
                  double* phi, int max_iter, double tol, double* residual, int* n_iter) {
    grid_t g;
    g.nr = nr; g.nth = nth; g.nz = nz;
    g.sy = nr + 4; g.sz = (long)(nr + 4) * (nth + 4); g.n = g.sz * (nz + 4);
    const double* c[7] = {aW, aE, aS, aN, aB, aT, aP};
    double* r = calloc((size_t)g.n, sizeof(double));
    double* z = calloc((size_t)g.n, sizeof(double));
    double* pv = calloc((size_t)g.n, sizeof(double));
    double* q = calloc((size_t)g.n, sizeof(double));
    if (!r || !z || !pv || !q) { free(r); free(z); free(pv); free(q); return -1; }

    matvec(&g, c, phi, q);
    axpy(&g, r, Su, -1.0, q);                      /* r = Su - A*phi */
    const double bnorm = sqrt(dot(&g, Su, Su));
    if (bnorm < SMALL) {
        *residual = 0.0;
        *n_iter = 0;
        free(r); free(z); free(pv); free(q);
        return 0;
    }
    jacobi(&g, z, r, aP);
    memcpy(pv, z, sizeof(double) * (size_t)g.n);
    double rz = dot(&g, r, z);
    *residual = sqrt(dot(&g, r, r)) / bnorm;
    *n_iter = 0;
    for (int iter = 1; iter <= max_iter; ++iter) {
        matvec(&g, c, pv, q);
        const double pq = dot(&g, pv, q);
        if (fabs(pq) < SMALL) break;
        const double alpha = rz / pq;
        axpy(&g, phi, phi, alpha, pv);
        axpy(&g, r, r, -alpha, q);
        *residual = sqrt(dot(&g, r, r)) / bnorm;
        *n_iter = iter;
        if (*residual < tol) break;
        jacobi(&g, z, r, aP);
        const double rz_new = dot(&g, r, z);
        const double beta = rz_new / fmax(rz, SMALL);
        rz = rz_new;
        axpy(&g, pv, z, beta, pv);
    }
    theta_halo(&g, phi);
    free(r); free(z); free(pv); free(q);
    return 0;
}
