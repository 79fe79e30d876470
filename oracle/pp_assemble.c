/*
 * oracle/pp_assemble.c -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement, plain C and FP64, of the assembly of EAF3D's p' system
 * (serial mesh).  It is the checker of the library's GPU assembler
 * (eafp_pp_assemble): only tests load it.
 *
 * Arrays use the reference's storage: Fortran order, bounds (-1:n+2) in
 * each axis, off(i,j,k) = (i+1) + (nr+4)*((j+1) + (nth+4)*(k+1)); the 1-D
 * coordinates r, theta, z have bounds (-1:n+2) (index + 1); the outlet mask
 * is the top layer, (nr+4)*(nth+4) in the same order.
 *
 * Restated from src/mod_pressure_3d.f90 (solve_pressure_correction):
 *   active phases            :127-169 (act_l from ws_liq_cont or alpha;
 *                                      act_g from alpha, ws_pv_active and
 *                                      the ws_gas_reach channel)
 *   zeroed system            :171-174
 *   cell gradients           :179-185 with cell_gradients :612-651,
 *                            pgrad (src/mod_constants.f90:274-287),
 *                            gas_gz :420-437, pv_cell :440-444; theta halos
 *                            of the gradients are the periodic copies
 *   phase contributions      :207-212 with add_phase_contribution
 *                                      :469-610, face_open :660-668,
 *                                      liq_weight_f :406-417
 *   gas compressibility      :226-246
 *   diagonal / trivial rows  :249-281 (PP_COMPLIANCE, P0_THERMO,
 *                                      C_SOUND_LIQ: :48, :63, :67)
 *   boundary conditions      src/mod_boundary_3d.f90:180-277
 *   continuity residual      :290-306
 *   liquid face exports      :525-530, :557-562, :597-602 (F* and a_nb of
 *                            the +faces, their link flags)
 * and, in oracle_pp_correct, the post-solve consumers:
 *   conservative fluxes      :333-351 (F = F* + a (pp_P - pp_nb), theta
 *                            halos by periodic copy)
 *   correct_velocities       :689-746, both phases
 *   pressure update / clamp  :358-360 (P_HYDRO_CAP :74)
 *   sealed cells             :373-397 (mean of the neighbours with fluid),
 *                            then the theta halo copy of p
 * The expressions keep the reference's operand order (left to right, as
 * written); built with -ffp-contract=off.
 *
 * Parity of this file is pinned by two test modules:
 *   tests/test_pp_oracle.py   face coefficients equal from both sides,
 *                             telescoping face fluxes (sum of Su over a
 *                             closed domain), the rest state (zero velocity
 *                             with a pressure gradient the cell and face
 *                             differences agree on gives Su = 0), the
 *                             diagonal bound and the boundary rows; for the
 *                             post-solve steps, balanced conservative fluxes,
 *                             the closed-form velocity change of a linear p'
 *                             and the clamped update / sealed cells;
 *   tests/test_oracle_pins.py the face coefficients of a single-phase mesh
 *                             against their closed form from the cell
 *                             geometry, the exact diagonal (compliance, gas
 *                             and liquid acoustic terms), and the gas
 *                             compressibility source of a uniform heating or
 *                             cooling step with its cap and T_MIN_GAS floor,
 *                             and the Rhie-Chow pressure term of u_f for a
 *                             pressure that alternates along r;
 *                             each check is also run on this file rebuilt
 *                             with a planted mistake (ORACLE_PLANT below) and
 *                             must fail there.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define SMALL 1.0e-30                 /* src/mod_constants.f90:53 */
#define ALPHA_FLOW_CUTOFF 1.0e-2      /* src/mod_constants.f90:68 */
#define GRAVITY 9.81                  /* src/mod_constants.f90:18 */
#define T_MIN_GAS 100.0               /* src/mod_constants.f90:148 */
#define PP_COMPLIANCE 1.0e-6          /* src/mod_pressure_3d.f90:48 */
#define COMP_SRC_CAP 0.2              /* src/mod_pressure_3d.f90:54 */
#define P0_THERMO 101325.0            /* src/mod_pressure_3d.f90:63 */
#define C_SOUND_LIQ 4000.0            /* src/mod_pressure_3d.f90:67 */
#define ROOM_EPS 1.0e-4               /* src/mod_pressure_3d.f90:663 */
#define P_HYDRO_CAP 2.0e6             /* src/mod_pressure_3d.f90:74 */

/* Mutation hooks of the pinning tests: tests/test_oracle_pins.py rebuilds
 * this file with -DORACLE_PLANT=n to show that each of its checks fails on
 * a plausible mistake.  Every other build leaves ORACLE_PLANT at 0, where
 * each hook is the restatement as written (same operations, same bits). */
#ifndef ORACLE_PLANT
#define ORACLE_PLANT 0
#endif
/* 1: harmonic instead of arithmetic mean of alpha V / aP_u on the r faces */
#if ORACLE_PLANT == 1
#define D_MEAN_R(a, b) (2.0 * (a) * (b) / ((a) + (b)))
#else
#define D_MEAN_R(a, b) (0.5 * ((a) + (b)))
#endif
/* 2: the own cell's alpha instead of the face average on the z faces */
#define ALPHA_MEAN_Z(a, b) (ORACLE_PLANT == 2 ? (a) : 0.5 * ((a) + (b)))
/* 3: the own cell's r area on the W face */
#define AREA_W(own, nb) (ORACLE_PLANT == 3 ? (own) : (nb))
/* 4: no compliance factor; 5: the gas acoustic term without 1/dt;
 * 6: c instead of c^2 in the liquid acoustic term */
#define COMPLIANCE (ORACLE_PLANT == 4 ? 0.0 : PP_COMPLIANCE)
#define ACOUSTIC_DT(dt) (ORACLE_PLANT == 5 ? 1.0 : (dt))
#define C2_LIQ (ORACLE_PLANT == 6 ? C_SOUND_LIQ : C_SOUND_LIQ * C_SOUND_LIQ)
/* 7: the compressibility source without its cap; 8: without the T_MIN_GAS floor */
#define COMP_CLAMP(cap, src) (ORACLE_PLANT == 7 ? (src) : fmax(-(cap), fmin((cap), (src))))
#define T_FLOOR (ORACLE_PLANT == 8 ? 0.0 : T_MIN_GAS)
/* 11: the Rhie-Chow term of the r faces' u_f without its d_f */
#define RC_D(d) (ORACLE_PLANT == 11 ? 1.0 : (d))

/* Same layout as eafp_pp_fields / eafp_pp_config in include/eafp.h. */
typedef struct {
    const double *r, *theta, *z;
    const double *vol, *Ar, *Ath, *Az;
    const int *cell_type;
    const int *outlet;
    const double *alpha_l, *ur_l, *uth_l, *uz_l, *aPr_l, *aPth_l, *aPz_l, *rho_l, *T_l;
    const double *alpha_g, *ur_g, *uth_g, *uz_g, *aPr_g, *aPth_g, *aPz_g, *rho_g, *T_old_g;
    const double *p;
    const int *liq_cont, *pv_active, *gas_reach;
    const double *liq_room;
} pp_fields;

typedef struct {
    double dt, rho_gas, T_ambient, beta_expansion;
    int gas_in_poisson, solve_multiphase, gas_compressibility, pad;
} pp_config;

typedef struct {
    int nr, nth, nz;
    long sy, sz, n;
} grid_t;

static inline long O(const grid_t* g, int i, int j, int k) {
    return (long)(i + 1) + g->sy * (j + 1) + g->sz * (k + 1);
}

static double pgrad(double pm, double p0, double pp_, double xm, double x0, double xp, int okm, int okp) {
    if (okm && okp) return (pp_ - pm) / (xp - xm);
    if (okp) return (pp_ - p0) / (xp - x0);
    if (okm) return (p0 - pm) / (x0 - xm);
    return 0.0;
}

typedef struct {
    const grid_t* g;
    const pp_fields* f;
    const pp_config* c;
    unsigned char *act_l, *act_g;
    double *gr, *gth, *gz, *gz_g;     /* liquid gradients; the gas phase differs in gz only */
    double *Fs[3], *ac[3];            /* liquid +face exports (r, theta, z), NULL: none */
    int* lk[3];
} ctx_t;

static int pv_cell(const ctx_t* x, long o) {
    return x->f->cell_type[o] != 0 && (x->act_l[o] || x->act_g[o]);
}

static double liq_weight_f(const ctx_t* x, int i, int j, int ka, int kb) {
    const pp_fields* f = x->f;
    const pp_config* c = x->c;
    const long a = O(x->g, i, j, ka), b = O(x->g, i, j, kb);
    return 0.5 * GRAVITY *
           ((x->act_l[a] ? 1.0 : 0.0) * f->alpha_l[a] * f->rho_l[a] * (1.0 - c->beta_expansion * (f->T_l[a] - c->T_ambient)) +
            (x->act_l[b] ? 1.0 : 0.0) * f->alpha_l[b] * f->rho_l[b] * (1.0 - c->beta_expansion * (f->T_l[b] - c->T_ambient)));
}

static void gradients(ctx_t* x) {
    const grid_t* g = x->g;
    const pp_fields* f = x->f;
    const double* p = f->p;
    for (int k = 1; k <= g->nz; ++k)
        for (int j = 1; j <= g->nth; ++j)
            for (int i = 1; i <= g->nr; ++i) {
                const long o = O(g, i, j, k);
                if (f->cell_type[o] == 0) continue;
                const long oW = O(g, i - 1, j, k), oE = O(g, i + 1, j, k);
                const long oS = O(g, i, j - 1, k), oN = O(g, i, j + 1, k);
                const long oB = O(g, i, j, k - 1), oT = O(g, i, j, k + 1);
                const double ri = f->r[i + 1];
                x->gr[o] = pgrad(p[oW], p[o], p[oE], f->r[i], f->r[i + 1], f->r[i + 2],
                                 i > 1 && pv_cell(x, oW), i < g->nr && pv_cell(x, oE));
                x->gth[o] = pgrad(p[oS], p[o], p[oN], ri * f->theta[j], ri * f->theta[j + 1], ri * f->theta[j + 2],
                                  pv_cell(x, oS), pv_cell(x, oN));
                const int okm = k > 1 && pv_cell(x, oB), okp = k < g->nz && pv_cell(x, oT);
                x->gz[o] = pgrad(p[oB], p[o], p[oT], f->z[k], f->z[k + 1], f->z[k + 2], okm, okp);
                double gb = 0.0, gt = 0.0;
                if (okm) gb = (p[o] - p[oB]) / (f->z[k + 1] - f->z[k]) + liq_weight_f(x, i, j, k - 1, k);
                if (okp) gt = (p[oT] - p[o]) / (f->z[k + 2] - f->z[k + 1]) + liq_weight_f(x, i, j, k, k + 1);
                x->gz_g[o] = okm && okp ? 0.5 * (gb + gt) : okp ? gt : okm ? gb : 0.0;
            }
    /* serial halo exchange: the theta planes 0 and nth+1 are periodic copies
     * (only these halo values of the gradients are read afterwards) */
    double* arrs[4] = {x->gr, x->gth, x->gz, x->gz_g};
    for (int a = 0; a < 4; ++a)
        for (int k = -1; k <= g->nz + 2; ++k)
            for (int i = -1; i <= g->nr + 2; ++i) {
                arrs[a][O(g, i, 0, k)] = arrs[a][O(g, i, g->nth, k)];
                arrs[a][O(g, i, -1, k)] = arrs[a][O(g, i, g->nth - 1, k)];
                arrs[a][O(g, i, g->nth + 1, k)] = arrs[a][O(g, i, 1, k)];
                arrs[a][O(g, i, g->nth + 2, k)] = arrs[a][O(g, i, 2, k)];
            }
}

static int face_open(const ctx_t* x, int isg, double u_f, long neg, long pos) {
    if (isg || !x->f->liq_room) return 1;
    if (u_f > 0.0 && x->f->liq_room[pos] <= ROOM_EPS) return 0;
    if (u_f < 0.0 && x->f->liq_room[neg] <= ROOM_EPS) return 0;
    return 1;
}

typedef struct {
    const double *alpha, *ur, *uth, *uz, *aPr, *aPth, *aPz;
    const unsigned char* act;
    const double* gz;
    int isg;
    int exp;        /* write the +face exports (the liquid) */
} phase_t;

static int link(const ctx_t* x, const unsigned char* act, long o) { return x->f->cell_type[o] != 0 && act[o]; }

static void add_phase(const ctx_t* x, const phase_t* ph, double* aW, double* aE, double* aS, double* aN, double* aB,
                      double* aT, double* Su, double* flux_ref) {
    const grid_t* g = x->g;
    const pp_fields* f = x->f;
    const double* pf = f->p;
    const double *gr = x->gr, *gth = x->gth, *gz = ph->gz;
    const double* aef = ph->alpha;
    for (int k = 1; k <= g->nz; ++k)
        for (int j = 1; j <= g->nth; ++j)
            for (int i = 1; i <= g->nr; ++i) {
                const long o = O(g, i, j, k);
                if (f->cell_type[o] == 0 || !ph->act[o]) continue;
                const long oW = O(g, i - 1, j, k), oE = O(g, i + 1, j, k);
                const long oS = O(g, i, j - 1, k), oN = O(g, i, j + 1, k);
                const long oB = O(g, i, j, k - 1), oT = O(g, i, j, k + 1);
                double d_f, af, u_f, delta;
                if (link(x, ph->act, oW)) {
                    d_f = D_MEAN_R(aef[o] * f->vol[o] / fmax(ph->aPr[o], SMALL), aef[oW] * f->vol[oW] / fmax(ph->aPr[oW], SMALL));
                    af = 0.5 * (aef[o] + aef[oW]);
                    delta = f->r[i + 1] - f->r[i];
                    u_f = 0.5 * (ph->ur[oW] + ph->ur[o]) + RC_D(d_f) * (0.5 * (gr[oW] + gr[o]) - (pf[o] - pf[oW]) / delta);
                    if (face_open(x, ph->isg, u_f, oW, o)) {
                        aW[o] = aW[o] + af * d_f * f->Ar[AREA_W(o, oW)] / delta;
                        Su[o] = Su[o] + af * u_f * f->Ar[oW];
                        *flux_ref = *flux_ref + fabs(af * u_f * f->Ar[oW]);
                    }
                }
                if (link(x, ph->act, oE)) {
                    d_f = D_MEAN_R(aef[o] * f->vol[o] / fmax(ph->aPr[o], SMALL), aef[oE] * f->vol[oE] / fmax(ph->aPr[oE], SMALL));
                    af = 0.5 * (aef[o] + aef[oE]);
                    delta = f->r[i + 2] - f->r[i + 1];
                    u_f = 0.5 * (ph->ur[o] + ph->ur[oE]) + RC_D(d_f) * (0.5 * (gr[o] + gr[oE]) - (pf[oE] - pf[o]) / delta);
                    if (face_open(x, ph->isg, u_f, o, oE)) {
                        aE[o] = aE[o] + af * d_f * f->Ar[o] / delta;
                        Su[o] = Su[o] - af * u_f * f->Ar[o];
                        *flux_ref = *flux_ref + fabs(af * u_f * f->Ar[o]);
                        if (ph->exp && x->Fs[0]) {
                            x->Fs[0][o] = af * u_f * f->Ar[o];
                            x->ac[0][o] = af * d_f * f->Ar[o] / delta;
                            x->lk[0][o] = 1;
                        }
                    }
                }
                if (link(x, ph->act, oS)) {
                    d_f = 0.5 * (aef[o] * f->vol[o] / fmax(ph->aPth[o], SMALL) + aef[oS] * f->vol[oS] / fmax(ph->aPth[oS], SMALL));
                    af = 0.5 * (aef[o] + aef[oS]);
                    delta = f->r[i + 1] * (f->theta[j + 1] - f->theta[j]);
                    u_f = 0.5 * (ph->uth[oS] + ph->uth[o]) + d_f * (0.5 * (gth[oS] + gth[o]) - (pf[o] - pf[oS]) / delta);
                    if (face_open(x, ph->isg, u_f, oS, o)) {
                        aS[o] = aS[o] + af * d_f * f->Ath[o] / delta;
                        Su[o] = Su[o] + af * u_f * f->Ath[o];
                        *flux_ref = *flux_ref + fabs(af * u_f * f->Ath[o]);
                    }
                }
                if (link(x, ph->act, oN)) {
                    d_f = 0.5 * (aef[o] * f->vol[o] / fmax(ph->aPth[o], SMALL) + aef[oN] * f->vol[oN] / fmax(ph->aPth[oN], SMALL));
                    af = 0.5 * (aef[o] + aef[oN]);
                    delta = f->r[i + 1] * (f->theta[j + 2] - f->theta[j + 1]);
                    u_f = 0.5 * (ph->uth[o] + ph->uth[oN]) + d_f * (0.5 * (gth[o] + gth[oN]) - (pf[oN] - pf[o]) / delta);
                    if (face_open(x, ph->isg, u_f, o, oN)) {
                        aN[o] = aN[o] + af * d_f * f->Ath[o] / delta;
                        Su[o] = Su[o] - af * u_f * f->Ath[o];
                        *flux_ref = *flux_ref + fabs(af * u_f * f->Ath[o]);
                        if (ph->exp && x->Fs[1]) {
                            x->Fs[1][o] = af * u_f * f->Ath[o];
                            x->ac[1][o] = af * d_f * f->Ath[o] / delta;
                            x->lk[1][o] = 1;
                        }
                    }
                }
                if (link(x, ph->act, oB)) {
                    d_f = 0.5 * (aef[o] * f->vol[o] / fmax(ph->aPz[o], SMALL) + aef[oB] * f->vol[oB] / fmax(ph->aPz[oB], SMALL));
                    af = ALPHA_MEAN_Z(aef[o], aef[oB]);
                    delta = f->z[k + 1] - f->z[k];
                    u_f = 0.5 * (ph->uz[oB] + ph->uz[o]) +
                          d_f * (0.5 * (gz[oB] + gz[o]) - (pf[o] - pf[oB]) / delta -
                                 (ph->isg ? liq_weight_f(x, i, j, k - 1, k) : 0.0));
                    if (face_open(x, ph->isg, u_f, oB, o)) {
                        aB[o] = aB[o] + af * d_f * f->Az[oB] / delta;
                        Su[o] = Su[o] + af * u_f * f->Az[oB];
                        *flux_ref = *flux_ref + fabs(af * u_f * f->Az[oB]);
                    }
                }
                if (link(x, ph->act, oT)) {
                    d_f = 0.5 * (aef[o] * f->vol[o] / fmax(ph->aPz[o], SMALL) + aef[oT] * f->vol[oT] / fmax(ph->aPz[oT], SMALL));
                    af = ALPHA_MEAN_Z(aef[o], aef[oT]);
                    delta = f->z[k + 2] - f->z[k + 1];
                    u_f = 0.5 * (ph->uz[o] + ph->uz[oT]) +
                          d_f * (0.5 * (gz[o] + gz[oT]) - (pf[oT] - pf[o]) / delta -
                                 (ph->isg ? liq_weight_f(x, i, j, k, k + 1) : 0.0));
                    if (face_open(x, ph->isg, u_f, o, oT)) {
                        aT[o] = aT[o] + af * d_f * f->Az[o] / delta;
                        Su[o] = Su[o] - af * u_f * f->Az[o];
                        *flux_ref = *flux_ref + fabs(af * u_f * f->Az[o]);
                        if (ph->exp && x->Fs[2]) {
                            x->Fs[2][o] = af * u_f * f->Az[o];
                            x->ac[2][o] = af * d_f * f->Az[o] / delta;
                            x->lk[2][o] = 1;
                        }
                    }
                }
            }
}

/* Liquid +face exports: F* and a_nb of the E, N, T faces and their link
 * flags (all zero where no open liquid face), padded arrays. */
typedef struct {
    double *Fs_r, *Fs_th, *Fs_z, *ac_r, *ac_th, *ac_z;
    int *lk_r, *lk_th, *lk_z;
} pp_exports;

/* Returns 0 on success, -1 if work arrays cannot be allocated.  Outputs are
 * the full padded arrays (halos zero), the continuity residual
 * sum|Su| / max(flux_ref, SMALL) with flux_ref, and (ex non-NULL) the liquid
 * face exports. */
int oracle_pp_assemble_ex(int nr, int nth, int nz, const pp_fields* f, const pp_config* c, double* aW, double* aE,
                          double* aS, double* aN, double* aB, double* aT, double* aP, double* Su, double* residual,
                          double* flux_ref_out, const pp_exports* ex) {
    grid_t g;
    g.nr = nr; g.nth = nth; g.nz = nz;
    g.sy = nr + 4; g.sz = (long)(nr + 4) * (nth + 4); g.n = g.sz * (nz + 4);
    ctx_t x;
    x.g = &g; x.f = f; x.c = c;
    x.Fs[0] = ex ? ex->Fs_r : NULL; x.Fs[1] = ex ? ex->Fs_th : NULL; x.Fs[2] = ex ? ex->Fs_z : NULL;
    x.ac[0] = ex ? ex->ac_r : NULL; x.ac[1] = ex ? ex->ac_th : NULL; x.ac[2] = ex ? ex->ac_z : NULL;
    x.lk[0] = ex ? ex->lk_r : NULL; x.lk[1] = ex ? ex->lk_th : NULL; x.lk[2] = ex ? ex->lk_z : NULL;
    for (int d = 0; d < 3; ++d)
        if (x.Fs[d]) {
            memset(x.Fs[d], 0, sizeof(double) * (size_t)g.n);
            memset(x.ac[d], 0, sizeof(double) * (size_t)g.n);
            memset(x.lk[d], 0, sizeof(int) * (size_t)g.n);
        }
    x.act_l = malloc((size_t)g.n);
    x.act_g = malloc((size_t)g.n);
    x.gr = calloc((size_t)g.n, sizeof(double));
    x.gth = calloc((size_t)g.n, sizeof(double));
    x.gz = calloc((size_t)g.n, sizeof(double));
    x.gz_g = calloc((size_t)g.n, sizeof(double));
    if (!x.act_l || !x.act_g || !x.gr || !x.gth || !x.gz || !x.gz_g) {
        free(x.act_l); free(x.act_g); free(x.gr); free(x.gth); free(x.gz); free(x.gz_g);
        return -1;
    }
    const int two_phase = c->gas_in_poisson && c->solve_multiphase;
    const int compress = c->gas_compressibility && c->solve_multiphase;
    for (long o = 0; o < g.n; ++o) {
        x.act_l[o] = f->liq_cont ? (f->liq_cont[o] != 0) : (f->alpha_l[o] >= ALPHA_FLOW_CUTOFF);
        int ag = f->alpha_g[o] >= ALPHA_FLOW_CUTOFF;
        if (f->pv_active) ag = ag && f->pv_active[o] != 0;
        if (f->gas_reach) ag = ag || (f->gas_reach[o] != 0 && f->alpha_g[o] < ALPHA_FLOW_CUTOFF && f->cell_type[o] != 0);
        x.act_g[o] = (unsigned char)ag;
    }
    double* outs[8] = {aW, aE, aS, aN, aB, aT, aP, Su};
    for (int a = 0; a < 8; ++a) memset(outs[a], 0, sizeof(double) * (size_t)g.n);
    gradients(&x);
    double flux_ref = 0.0;
    phase_t liq = {f->alpha_l, f->ur_l, f->uth_l, f->uz_l, f->aPr_l, f->aPth_l, f->aPz_l, x.act_l, x.gz, 0, 1};
    add_phase(&x, &liq, aW, aE, aS, aN, aB, aT, Su, &flux_ref);
    if (two_phase) {
        phase_t gas = {f->alpha_g, f->ur_g, f->uth_g, f->uz_g, f->aPr_g, f->aPth_g, f->aPz_g, x.act_g, x.gz_g, 1, 0};
        add_phase(&x, &gas, aW, aE, aS, aN, aB, aT, Su, &flux_ref);
    }
    if (compress) {
        for (int k = 1; k <= nz; ++k)
            for (int j = 1; j <= nth; ++j)
                for (int i = 1; i <= nr; ++i) {
                    const long o = O(&g, i, j, k);
                    if (f->cell_type[o] == 0 || !x.act_g[o]) continue;
                    const double src = f->alpha_g[o] * f->vol[o] / c->dt *
                                       (f->rho_g[o] - c->rho_gas * c->T_ambient / fmax(f->T_old_g[o], T_FLOOR)) /
                                       fmax(f->rho_g[o], SMALL);
                    const double cap = COMP_SRC_CAP * f->alpha_g[o] * f->vol[o] / c->dt;
                    Su[o] = Su[o] - COMP_CLAMP(cap, src);
                }
    }
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= nth; ++j)
            for (int i = 1; i <= nr; ++i) {
                const long o = O(&g, i, j, k);
                if (f->cell_type[o] == 0) continue;
                if (!x.act_l[o] && !x.act_g[o]) {
                    aW[o] = 0.0; aE[o] = 0.0; aS[o] = 0.0; aN[o] = 0.0; aB[o] = 0.0; aT[o] = 0.0;
                    aP[o] = 1.0; Su[o] = 0.0;
                } else {
                    aP[o] = (aW[o] + aE[o] + aS[o] + aN[o] + aB[o] + aT[o]) * (1.0 + COMPLIANCE);
                    if (compress && x.act_g[o]) aP[o] = aP[o] + f->alpha_g[o] / P0_THERMO * f->vol[o] / ACOUSTIC_DT(c->dt);
                    if (x.act_l[o])
                        aP[o] = aP[o] + f->alpha_l[o] * f->vol[o] / (f->rho_l[o] * C2_LIQ * c->dt);
                }
            }
    /* src/mod_boundary_3d.f90:180-277 (serial: every boundary is physical) */
    const long sy2 = nr + 4;
#define OUT(i, j) (f->outlet[(long)((i) + 1) + sy2 * ((j) + 1)] != 0)
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= nth; ++j)
            for (int i = 1; i <= nr; ++i) {
                const long o = O(&g, i, j, k);
                if (f->cell_type[o] == 0) {
                    aW[o] = 0.0; aE[o] = 0.0; aS[o] = 0.0; aN[o] = 0.0; aB[o] = 0.0; aT[o] = 0.0;
                    aP[o] = 1.0; Su[o] = 0.0;
                    continue;
                }
                if (i == 1) aW[o] = 0.0;
                if (i == nr) aE[o] = 0.0;
                if (k == 1) aB[o] = 0.0;
                if (k == nz) {
                    if (OUT(i, j)) {
                        aW[o] = 0.0; aE[o] = 0.0; aS[o] = 0.0; aN[o] = 0.0; aB[o] = 0.0; aT[o] = 0.0;
                        aP[o] = 1.0; Su[o] = 0.0;
                        continue;
                    }
                    aT[o] = 0.0;
                }
                if (i > 1 && f->cell_type[O(&g, i - 1, j, k)] == 0) aW[o] = 0.0;
                if (i < nr && f->cell_type[O(&g, i + 1, j, k)] == 0) aE[o] = 0.0;
                if (k > 1 && f->cell_type[O(&g, i, j, k - 1)] == 0) aB[o] = 0.0;
                if (k < nz && f->cell_type[O(&g, i, j, k + 1)] == 0) aT[o] = 0.0;
                if (k == nz) {          /* the outlet mask lives on the top layer only */
                    if (OUT(i - 1, j)) aW[o] = 0.0;
                    if (OUT(i + 1, j)) aE[o] = 0.0;
                    if (OUT(i, j - 1)) aS[o] = 0.0;
                    if (OUT(i, j + 1)) aN[o] = 0.0;
                }
                if (k + 1 == nz && OUT(i, j)) aT[o] = 0.0;
            }
#undef OUT
    double sum_su = 0.0;
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= nth; ++j)
            for (int i = 1; i <= nr; ++i) {
                const long o = O(&g, i, j, k);
                if (f->cell_type[o] == 0) continue;
                if (x.act_l[o] || x.act_g[o]) sum_su = sum_su + fabs(Su[o]);
            }
    *residual = sum_su / fmax(flux_ref, SMALL);
    *flux_ref_out = flux_ref;
    free(x.act_l); free(x.act_g); free(x.gr); free(x.gth); free(x.gz); free(x.gz_g);
    return 0;
}

int oracle_pp_assemble(int nr, int nth, int nz, const pp_fields* f, const pp_config* c, double* aW, double* aE,
                       double* aS, double* aN, double* aB, double* aT, double* aP, double* Su, double* residual,
                       double* flux_ref_out) {
    return oracle_pp_assemble_ex(nr, nth, nz, f, c, aW, aE, aS, aN, aB, aT, aP, Su, residual, flux_ref_out, NULL);
}

static void theta_copy_d(const grid_t* g, double* a) {
    const int dst[4] = {0, -1, g->nth + 1, g->nth + 2};
    const int src[4] = {g->nth, g->nth - 1, 1, 2};
    for (int q = 0; q < 4; ++q)
        for (int k = -1; k <= g->nz + 2; ++k)
            for (int i = -1; i <= g->nr + 2; ++i) a[O(g, i, dst[q], k)] = a[O(g, i, src[q], k)];
}

static void theta_copy_i(const grid_t* g, int* a) {
    const int dst[4] = {0, -1, g->nth + 1, g->nth + 2};
    const int src[4] = {g->nth, g->nth - 1, 1, 2};
    for (int q = 0; q < 4; ++q)
        for (int k = -1; k <= g->nz + 2; ++k)
            for (int i = -1; i <= g->nr + 2; ++i) a[O(g, i, dst[q], k)] = a[O(g, i, src[q], k)];
}

/* Post-solve consumers (see the file header).  pp: the solved correction
 * (padded, theta halos filled by the CG).  ex: the assembly's exports;
 * Fc/lk_out: conservative fluxes and link flags (padded, zero where no open
 * liquid face, theta halos by periodic copy).  vel: ur, uth, uz of the
 * liquid then of the gas, corrected in place; p updated in place. */
int oracle_pp_correct(int nr, int nth, int nz, const pp_fields* f, const pp_config* c, double alpha_p, const double* pp,
                      const pp_exports* ex, double* Fc_r, double* Fc_th, double* Fc_z, int* lk_r, int* lk_th,
                      int* lk_z, double* const vel[6], double* p) {
    grid_t g;
    g.nr = nr; g.nth = nth; g.nz = nz;
    g.sy = nr + 4; g.sz = (long)(nr + 4) * (nth + 4); g.n = g.sz * (nz + 4);
    unsigned char* act_l = malloc((size_t)g.n);
    unsigned char* act_g = malloc((size_t)g.n);
    if (!act_l || !act_g) { free(act_l); free(act_g); return -1; }
    for (long o = 0; o < g.n; ++o) {
        act_l[o] = f->liq_cont ? (f->liq_cont[o] != 0) : (f->alpha_l[o] >= ALPHA_FLOW_CUTOFF);
        int ag = f->alpha_g[o] >= ALPHA_FLOW_CUTOFF;
        if (f->pv_active) ag = ag && f->pv_active[o] != 0;
        if (f->gas_reach) ag = ag || (f->gas_reach[o] != 0 && f->alpha_g[o] < ALPHA_FLOW_CUTOFF && f->cell_type[o] != 0);
        act_g[o] = (unsigned char)ag;
    }
    /* conservative liquid fluxes (:333-351) */
    memset(Fc_r, 0, sizeof(double) * (size_t)g.n);
    memset(Fc_th, 0, sizeof(double) * (size_t)g.n);
    memset(Fc_z, 0, sizeof(double) * (size_t)g.n);
    memcpy(lk_r, ex->lk_r, sizeof(int) * (size_t)g.n);
    memcpy(lk_th, ex->lk_th, sizeof(int) * (size_t)g.n);
    memcpy(lk_z, ex->lk_z, sizeof(int) * (size_t)g.n);
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= nth; ++j)
            for (int i = 1; i <= nr; ++i) {
                const long o = O(&g, i, j, k);
                if (f->cell_type[o] == 0) continue;
                if (ex->lk_r[o] == 1) Fc_r[o] = ex->Fs_r[o] + ex->ac_r[o] * (pp[o] - pp[O(&g, i + 1, j, k)]);
                if (ex->lk_th[o] == 1) Fc_th[o] = ex->Fs_th[o] + ex->ac_th[o] * (pp[o] - pp[O(&g, i, j + 1, k)]);
                if (ex->lk_z[o] == 1) Fc_z[o] = ex->Fs_z[o] + ex->ac_z[o] * (pp[o] - pp[O(&g, i, j, k + 1)]);
            }
    theta_copy_d(&g, Fc_r); theta_copy_d(&g, Fc_th); theta_copy_d(&g, Fc_z);
    theta_copy_i(&g, lk_r); theta_copy_i(&g, lk_th); theta_copy_i(&g, lk_z);
    /* correct_velocities (:689-746): liquid, then the gas with the two-phase coupling */
    const int two_phase = c->gas_in_poisson && c->solve_multiphase;
    for (int ph = 0; ph < (two_phase ? 2 : 1); ++ph) {
        const unsigned char* act = ph ? act_g : act_l;
        const double* alpha = ph ? f->alpha_g : f->alpha_l;
        const double* aPr = ph ? f->aPr_g : f->aPr_l;
        const double* aPth = ph ? f->aPth_g : f->aPth_l;
        const double* aPz = ph ? f->aPz_g : f->aPz_l;
        double *ur = vel[3 * ph], *uth = vel[3 * ph + 1], *uz = vel[3 * ph + 2];
        for (int k = 1; k <= nz; ++k)
            for (int j = 1; j <= nth; ++j)
                for (int i = 1; i <= nr; ++i) {
                    const long o = O(&g, i, j, k);
                    if (f->cell_type[o] == 0 || !act[o]) continue;
                    double d_coeff;
                    if (!(i == 1 || i == nr) && fabs(aPr[o]) > SMALL) {
                        d_coeff = alpha[o] * f->vol[o] / aPr[o];
                        ur[o] = ur[o] - d_coeff * (pp[O(&g, i + 1, j, k)] - pp[O(&g, i - 1, j, k)]) / (f->r[i + 2] - f->r[i]);
                    }
                    if (fabs(aPth[o]) > SMALL) {
                        d_coeff = alpha[o] * f->vol[o] / aPth[o];
                        uth[o] = uth[o] - d_coeff * (pp[O(&g, i, j + 1, k)] - pp[O(&g, i, j - 1, k)]) /
                                              (f->r[i + 1] * (f->theta[j + 2] - f->theta[j]));
                    }
                    if (!(k == 1 || k == nz) && fabs(aPz[o]) > SMALL) {
                        d_coeff = alpha[o] * f->vol[o] / aPz[o];
                        uz[o] = uz[o] - d_coeff * (pp[O(&g, i, j, k + 1)] - pp[O(&g, i, j, k - 1)]) / (f->z[k + 2] - f->z[k]);
                    }
                }
    }
    /* pressure update and clamp over the whole array (:358-360) */
    for (long o = 0; o < g.n; ++o) {
        p[o] = p[o] + alpha_p * pp[o];
        p[o] = fmax(-P_HYDRO_CAP, fmin(P_HYDRO_CAP, p[o]));
    }
    /* sealed cells: mean of the neighbours with fluid (:373-397) */
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= nth; ++j)
            for (int i = 1; i <= nr; ++i) {
                const long o = O(&g, i, j, k);
                if (f->cell_type[o] == 0 || act_l[o] || act_g[o]) continue;
                const long nb[6] = {O(&g, i - 1, j, k), O(&g, i + 1, j, k), O(&g, i, j - 1, k),
                                    O(&g, i, j + 1, k), O(&g, i, j, k - 1), O(&g, i, j, k + 1)};
                double psum = 0.0;
                int nnb = 0;
                for (int q = 0; q < 6; ++q) {
                    const long b = nb[q];
                    if (f->cell_type[b] == 0 || (!act_l[b] && !act_g[b])) continue;
                    psum = psum + p[b];
                    nnb = nnb + 1;
                }
                p[o] = nnb > 0 ? psum / (double)nnb : 0.0;
            }
    theta_copy_d(&g, p);
    free(act_l); free(act_g);
    return 0;
}
