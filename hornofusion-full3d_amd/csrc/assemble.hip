// assemble.hip -- the p' system of EAF3D assembled on the GPU from the phase
// fields: what solve_pressure_correction does before its cg_3d_mpi call
// (src/mod_pressure_3d.f90:127-306 with add_phase_contribution :469-610,
// cell_gradients :612-651, and apply_pressure_bc,
// src/mod_boundary_3d.f90:180-277), for one serial mesh.
//
// Layout: the caller's padded Fortran arrays (bounds (-1:n+2), r fastest);
// one thread per padded cell, halo cells write zeros.  Two kernels:
//   k_pp_grad      cell pressure gradients gr, gth, gz of the liquid and gz
//                  of the gas (whose vertical faces drop the liquid weight);
//                  the theta halos of the gradients are read through the
//                  periodic index, which is what the reference's serial halo
//                  copy leaves there
//   k_pp_assemble  per cell: the Rhie-Chow face terms of the liquid, then
//                  of the gas, the gas compressibility source, the diagonal
//                  (compliance, acoustic terms) or the trivial row, the
//                  boundary rows (walls, outlet Dirichlet and its folding),
//                  and per-block partial sums of |face flux| and |Su|
// Every expression keeps the reference's operand order and the file is
// compiled without floating-point contraction, so each coefficient is the
// same double the oracle restatement computes.  The continuity residual's
// two global sums are reduced in block order on the host.
//
// Algorithmic bytes per padded cell: the two gradient passes read p,
// cell_type and the activity inputs of the cell and its 6 neighbours
// (cached), write 4 gradients (32 B); the assembly reads ~26 input arrays
// (~200 B) and the gradients of 7 cells, writes 8 coefficient arrays (64 B).
#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "batch.hpp"

namespace eafp {

namespace {

constexpr double ALPHA_FLOW_CUTOFF = 1.0e-2;   // src/mod_constants.f90:68
constexpr double GRAVITY = 9.81;               // src/mod_constants.f90:18
constexpr double T_MIN_GAS = 100.0;            // src/mod_constants.f90:148
constexpr double PP_COMPLIANCE = 1.0e-6;       // src/mod_pressure_3d.f90:48
constexpr double COMP_SRC_CAP = 0.2;           // src/mod_pressure_3d.f90:54
constexpr double P0_THERMO = 101325.0;         // src/mod_pressure_3d.f90:63
constexpr double C_SOUND_LIQ = 4000.0;         // src/mod_pressure_3d.f90:67
constexpr double ROOM_EPS = 1.0e-4;            // src/mod_pressure_3d.f90:663
constexpr int NF3 = 23;                        // 3-D double inputs (vol .. p)

struct PPArgs {
    int nr, nth, nz;
    long sy, sz, n;          // padded strides and length
    const double *r, *th, *z;
    const double *vol, *Ar, *Ath, *Az;
    const int* ct;
    const int* outlet;
    const double *al, *url, *uthl, *uzl, *aPrl, *aPthl, *aPzl, *rhol, *Tl;
    const double *ag, *urg, *uthg, *uzg, *aPrg, *aPthg, *aPzg, *rhog, *Toldg;
    const double* p;
    const int *liq_cont, *pv_active, *gas_reach;
    const double* liq_room;
    double dt, rho_gas, T_amb, beta;
    int two_phase, compress;
    double *gr, *gth, *gz, *gzg;
    double *aW, *aE, *aS, *aN, *aB, *aT, *aP, *Su;
    double* part;            // 2 per block: flux_ref, sum|Su|
};

__device__ __forceinline__ bool act_l(const PPArgs& a, long o) {
    return a.liq_cont ? a.liq_cont[o] != 0 : a.al[o] >= ALPHA_FLOW_CUTOFF;
}

__device__ __forceinline__ bool act_g(const PPArgs& a, long o) {
    bool g = a.ag[o] >= ALPHA_FLOW_CUTOFF;
    if (a.pv_active) g = g && a.pv_active[o] != 0;
    if (a.gas_reach) g = g || (a.gas_reach[o] != 0 && a.ag[o] < ALPHA_FLOW_CUTOFF && a.ct[o] != 0);
    return g;
}

__device__ __forceinline__ bool pv_cell(const PPArgs& a, long o) { return a.ct[o] != 0 && (act_l(a, o) || act_g(a, o)); }

__device__ __forceinline__ double pgrad(double pm, double p0, double pp_, double xm, double x0, double xp, bool okm,
                                        bool okp) {
    if (okm && okp) return (pp_ - pm) / (xp - xm);
    if (okp) return (pp_ - p0) / (xp - x0);
    if (okm) return (p0 - pm) / (x0 - xm);
    return 0.0;
}

// Liquid weight on the z-face between cells a and b (src/mod_pressure_3d.f90:406-417).
__device__ __forceinline__ double liq_weight(const PPArgs& x, long a, long b) {
    return 0.5 * GRAVITY *
           ((act_l(x, a) ? 1.0 : 0.0) * x.al[a] * x.rhol[a] * (1.0 - x.beta * (x.Tl[a] - x.T_amb)) +
            (act_l(x, b) ? 1.0 : 0.0) * x.al[b] * x.rhol[b] * (1.0 - x.beta * (x.Tl[b] - x.T_amb)));
}

struct Pos {
    int i, j, k;
};

__device__ __forceinline__ Pos padded_pos(const PPArgs& a, long o) {
    const int k = (int)(o / a.sz);
    const long rem = o - (long)k * a.sz;
    const int j = (int)(rem / a.sy);
    const int i = (int)(rem - (long)j * a.sy);
    return {i - 1, j - 1, k - 1};
}

__device__ __forceinline__ long at(const PPArgs& a, int i, int j, int k) {
    return (long)(i + 1) + a.sy * (j + 1) + a.sz * (k + 1);
}

__global__ __launch_bounds__(BLOCK) void k_pp_grad(PPArgs a) {
    const long o = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (o >= a.n) return;
    const Pos q = padded_pos(a, o);
    const int i = q.i, j = q.j, k = q.k;
    const bool interior = i >= 1 && i <= a.nr && j >= 1 && j <= a.nth && k >= 1 && k <= a.nz;
    double gr = 0.0, gth = 0.0, gz = 0.0, gzg = 0.0;
    if (interior && a.ct[o] != 0) {
        const long oW = at(a, i - 1, j, k), oE = at(a, i + 1, j, k);
        const long oS = at(a, i, j - 1, k), oN = at(a, i, j + 1, k);
        const long oB = at(a, i, j, k - 1), oT = at(a, i, j, k + 1);
        const double* p = a.p;
        const double ri = a.r[i + 1];
        gr = pgrad(p[oW], p[o], p[oE], a.r[i], a.r[i + 1], a.r[i + 2], i > 1 && pv_cell(a, oW),
                   i < a.nr && pv_cell(a, oE));
        gth = pgrad(p[oS], p[o], p[oN], ri * a.th[j], ri * a.th[j + 1], ri * a.th[j + 2], pv_cell(a, oS),
                    pv_cell(a, oN));
        const bool okm = k > 1 && pv_cell(a, oB), okp = k < a.nz && pv_cell(a, oT);
        gz = pgrad(p[oB], p[o], p[oT], a.z[k], a.z[k + 1], a.z[k + 2], okm, okp);
        double gb = 0.0, gt = 0.0;
        if (okm) gb = (p[o] - p[oB]) / (a.z[k + 1] - a.z[k]) + liq_weight(a, oB, o);
        if (okp) gt = (p[oT] - p[o]) / (a.z[k + 2] - a.z[k + 1]) + liq_weight(a, o, oT);
        gzg = okm && okp ? 0.5 * (gb + gt) : okp ? gt : okm ? gb : 0.0;
    }
    a.gr[o] = gr;
    a.gth[o] = gth;
    a.gz[o] = gz;
    a.gzg[o] = gzg;
}

// One phase's view for the face terms.
struct Phase {
    const double *alpha, *ur, *uth, *uz, *aPr, *aPth, *aPz;
    const double* gz;
    bool gas;
};

__device__ __forceinline__ bool face_open(const PPArgs& a, bool isg, double u_f, long neg, long pos) {
    if (isg || !a.liq_room) return true;
    if (u_f > 0.0 && a.liq_room[pos] <= ROOM_EPS) return false;
    if (u_f < 0.0 && a.liq_room[neg] <= ROOM_EPS) return false;
    return true;
}

struct Acc {
    double aW, aE, aS, aN, aB, aT, Su, flux;
};

// The six faces of cell o for one phase (src/mod_pressure_3d.f90:486-607).
__device__ void add_phase(const PPArgs& a, const Phase& ph, int i, int j, int k, long o, Acc& c) {
    const bool gas = ph.gas;
    auto act = [&](long q) { return gas ? act_g(a, q) : act_l(a, q); };
    auto link = [&](long q) { return a.ct[q] != 0 && act(q); };
    const double* aef = ph.alpha;
    const double* pf = a.p;
    const long oW = at(a, i - 1, j, k), oE = at(a, i + 1, j, k);
    const long oS = at(a, i, j - 1, k), oN = at(a, i, j + 1, k);
    const long oB = at(a, i, j, k - 1), oT = at(a, i, j, k + 1);
    // theta halos of the gradients: the periodic copies
    const int js = j - 1 == 0 ? a.nth : j - 1, jn = j + 1 == a.nth + 1 ? 1 : j + 1;
    const long gS = at(a, i, js, k), gN = at(a, i, jn, k);
    const double* gr = a.gr;
    const double* gth = a.gth;
    const double* gz = ph.gz;
    double d_f, af, u_f, delta;
    if (link(oW)) {
        d_f = 0.5 * (aef[o] * a.vol[o] / fmax(ph.aPr[o], SMALL) + aef[oW] * a.vol[oW] / fmax(ph.aPr[oW], SMALL));
        af = 0.5 * (aef[o] + aef[oW]);
        delta = a.r[i + 1] - a.r[i];
        u_f = 0.5 * (ph.ur[oW] + ph.ur[o]) + d_f * (0.5 * (gr[oW] + gr[o]) - (pf[o] - pf[oW]) / delta);
        if (face_open(a, gas, u_f, oW, o)) {
            c.aW = c.aW + af * d_f * a.Ar[oW] / delta;
            c.Su = c.Su + af * u_f * a.Ar[oW];
            c.flux += fabs(af * u_f * a.Ar[oW]);
        }
    }
    if (link(oE)) {
        d_f = 0.5 * (aef[o] * a.vol[o] / fmax(ph.aPr[o], SMALL) + aef[oE] * a.vol[oE] / fmax(ph.aPr[oE], SMALL));
        af = 0.5 * (aef[o] + aef[oE]);
        delta = a.r[i + 2] - a.r[i + 1];
        u_f = 0.5 * (ph.ur[o] + ph.ur[oE]) + d_f * (0.5 * (gr[o] + gr[oE]) - (pf[oE] - pf[o]) / delta);
        if (face_open(a, gas, u_f, o, oE)) {
            c.aE = c.aE + af * d_f * a.Ar[o] / delta;
            c.Su = c.Su - af * u_f * a.Ar[o];
            c.flux += fabs(af * u_f * a.Ar[o]);
        }
    }
    if (link(oS)) {
        d_f = 0.5 * (aef[o] * a.vol[o] / fmax(ph.aPth[o], SMALL) + aef[oS] * a.vol[oS] / fmax(ph.aPth[oS], SMALL));
        af = 0.5 * (aef[o] + aef[oS]);
        delta = a.r[i + 1] * (a.th[j + 1] - a.th[j]);
        u_f = 0.5 * (ph.uth[oS] + ph.uth[o]) + d_f * (0.5 * (gth[gS] + gth[o]) - (pf[o] - pf[oS]) / delta);
        if (face_open(a, gas, u_f, oS, o)) {
            c.aS = c.aS + af * d_f * a.Ath[o] / delta;
            c.Su = c.Su + af * u_f * a.Ath[o];
            c.flux += fabs(af * u_f * a.Ath[o]);
        }
    }
    if (link(oN)) {
        d_f = 0.5 * (aef[o] * a.vol[o] / fmax(ph.aPth[o], SMALL) + aef[oN] * a.vol[oN] / fmax(ph.aPth[oN], SMALL));
        af = 0.5 * (aef[o] + aef[oN]);
        delta = a.r[i + 1] * (a.th[j + 2] - a.th[j + 1]);
        u_f = 0.5 * (ph.uth[o] + ph.uth[oN]) + d_f * (0.5 * (gth[o] + gth[gN]) - (pf[oN] - pf[o]) / delta);
        if (face_open(a, gas, u_f, o, oN)) {
            c.aN = c.aN + af * d_f * a.Ath[o] / delta;
            c.Su = c.Su - af * u_f * a.Ath[o];
            c.flux += fabs(af * u_f * a.Ath[o]);
        }
    }
    if (link(oB)) {
        d_f = 0.5 * (aef[o] * a.vol[o] / fmax(ph.aPz[o], SMALL) + aef[oB] * a.vol[oB] / fmax(ph.aPz[oB], SMALL));
        af = 0.5 * (aef[o] + aef[oB]);
        delta = a.z[k + 1] - a.z[k];
        u_f = 0.5 * (ph.uz[oB] + ph.uz[o]) +
              d_f * (0.5 * (gz[oB] + gz[o]) - (pf[o] - pf[oB]) / delta - (gas ? liq_weight(a, oB, o) : 0.0));
        if (face_open(a, gas, u_f, oB, o)) {
            c.aB = c.aB + af * d_f * a.Az[oB] / delta;
            c.Su = c.Su + af * u_f * a.Az[oB];
            c.flux += fabs(af * u_f * a.Az[oB]);
        }
    }
    if (link(oT)) {
        d_f = 0.5 * (aef[o] * a.vol[o] / fmax(ph.aPz[o], SMALL) + aef[oT] * a.vol[oT] / fmax(ph.aPz[oT], SMALL));
        af = 0.5 * (aef[o] + aef[oT]);
        delta = a.z[k + 2] - a.z[k + 1];
        u_f = 0.5 * (ph.uz[o] + ph.uz[oT]) +
              d_f * (0.5 * (gz[o] + gz[oT]) - (pf[oT] - pf[o]) / delta - (gas ? liq_weight(a, o, oT) : 0.0));
        if (face_open(a, gas, u_f, o, oT)) {
            c.aT = c.aT + af * d_f * a.Az[o] / delta;
            c.Su = c.Su - af * u_f * a.Az[o];
            c.flux += fabs(af * u_f * a.Az[o]);
        }
    }
}

__global__ __launch_bounds__(BLOCK) void k_pp_assemble(PPArgs a) {
    __shared__ double sh[2 * (BLOCK / 64)];
    const long o = (long)blockIdx.x * BLOCK + threadIdx.x;
    double acc[2] = {0.0, 0.0};
    if (o < a.n) {
        const Pos q = padded_pos(a, o);
        const int i = q.i, j = q.j, k = q.k;
        const bool interior = i >= 1 && i <= a.nr && j >= 1 && j <= a.nth && k >= 1 && k <= a.nz;
        Acc c = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        double aP = 0.0;
        if (interior && a.ct[o] != 0) {
            const bool al = act_l(a, o), ag = act_g(a, o);
            if (al) {
                const Phase L = {a.al, a.url, a.uthl, a.uzl, a.aPrl, a.aPthl, a.aPzl, a.gz, false};
                add_phase(a, L, i, j, k, o, c);
            }
            if (a.two_phase && ag) {
                const Phase Gp = {a.ag, a.urg, a.uthg, a.uzg, a.aPrg, a.aPthg, a.aPzg, a.gzg, true};
                add_phase(a, Gp, i, j, k, o, c);
            }
            if (a.compress && ag) {
                const double src = a.ag[o] * a.vol[o] / a.dt *
                                   (a.rhog[o] - a.rho_gas * a.T_amb / fmax(a.Toldg[o], T_MIN_GAS)) / fmax(a.rhog[o], SMALL);
                const double cap = COMP_SRC_CAP * a.ag[o] * a.vol[o] / a.dt;
                c.Su = c.Su - fmax(-cap, fmin(cap, src));
            }
            if (!al && !ag) {
                c.aW = c.aE = c.aS = c.aN = c.aB = c.aT = 0.0;
                aP = 1.0;
                c.Su = 0.0;
            } else {
                aP = (c.aW + c.aE + c.aS + c.aN + c.aB + c.aT) * (1.0 + PP_COMPLIANCE);
                if (a.compress && ag) aP = aP + a.ag[o] / P0_THERMO * a.vol[o] / a.dt;
                if (al) aP = aP + a.al[o] * a.vol[o] / (a.rhol[o] * (C_SOUND_LIQ * C_SOUND_LIQ) * a.dt);
            }
            // boundary rows (src/mod_boundary_3d.f90:218-275; serial: every boundary physical)
            const long sy2 = a.nr + 4;
            auto outlet = [&](int ii, int jj) { return a.outlet[(long)(ii + 1) + sy2 * (jj + 1)] != 0; };
            bool ident = false;
            if (i == 1) c.aW = 0.0;
            if (i == a.nr) c.aE = 0.0;
            if (k == 1) c.aB = 0.0;
            if (k == a.nz) {
                if (outlet(i, j)) ident = true;
                else c.aT = 0.0;
            }
            if (ident) {
                c.aW = c.aE = c.aS = c.aN = c.aB = c.aT = 0.0;
                aP = 1.0;
                c.Su = 0.0;
            } else {
                if (i > 1 && a.ct[at(a, i - 1, j, k)] == 0) c.aW = 0.0;
                if (i < a.nr && a.ct[at(a, i + 1, j, k)] == 0) c.aE = 0.0;
                if (k > 1 && a.ct[at(a, i, j, k - 1)] == 0) c.aB = 0.0;
                if (k < a.nz && a.ct[at(a, i, j, k + 1)] == 0) c.aT = 0.0;
                if (k == a.nz) {
                    if (outlet(i - 1, j)) c.aW = 0.0;
                    if (outlet(i + 1, j)) c.aE = 0.0;
                    if (outlet(i, j - 1)) c.aS = 0.0;
                    if (outlet(i, j + 1)) c.aN = 0.0;
                }
                if (k + 1 == a.nz && outlet(i, j)) c.aT = 0.0;
            }
            acc[0] = c.flux;
            if (al || ag) acc[1] = fabs(c.Su);
        } else if (interior) {
            aP = 1.0;                        // masked cell: identity row
        }
        a.aW[o] = c.aW;
        a.aE[o] = c.aE;
        a.aS[o] = c.aS;
        a.aN[o] = c.aN;
        a.aB[o] = c.aB;
        a.aT[o] = c.aT;
        a.aP[o] = aP;
        a.Su[o] = c.Su;
    }
    block_sum<2>(acc, sh);
    if (threadIdx.x == 0) {
        a.part[2 * (long)blockIdx.x] = acc[0];
        a.part[2 * (long)blockIdx.x + 1] = acc[1];
    }
}

struct PPWork {
    int nr = -1, nth = -1, nz = -1;
    hipStream_t stream = nullptr;
    DevBuf<double> coord, f3, grads, outs, part, room;
    DevBuf<int> ct, outlet, masks;
    std::vector<double> part_h;
};

std::mutex g_pp_mutex;
PPWork* g_pp = nullptr;

int ensure_pp(const eafp_desc* d) {
    if (!g_pp) g_pp = new PPWork();
    PPWork& w = *g_pp;
    if (!w.stream) EAFP_TRY(hipStreamCreateWithFlags(&w.stream, hipStreamNonBlocking));
    if (w.nr == d->nr && w.nth == d->ntheta && w.nz == d->nz) return EAFP_OK;
    const long n = (long)(d->nr + 4) * (d->ntheta + 4) * (d->nz + 4);
    EAFP_TRY_RC(w.coord.alloc((size_t)(d->nr + d->ntheta + d->nz + 12)));
    EAFP_TRY_RC(w.f3.alloc((size_t)n * NF3));
    EAFP_TRY_RC(w.grads.alloc((size_t)n * 4));
    EAFP_TRY_RC(w.outs.alloc((size_t)n * 8));
    EAFP_TRY_RC(w.room.alloc((size_t)n));
    EAFP_TRY_RC(w.ct.alloc((size_t)n));
    EAFP_TRY_RC(w.masks.alloc((size_t)n * 3));
    EAFP_TRY_RC(w.outlet.alloc((size_t)(d->nr + 4) * (d->ntheta + 4)));
    const long blocks = (n + BLOCK - 1) / BLOCK;
    EAFP_TRY_RC(w.part.alloc((size_t)blocks * 2));
    w.part_h.assign((size_t)blocks * 2, 0.0);
    w.nr = d->nr;
    w.nth = d->ntheta;
    w.nz = d->nz;
    return EAFP_OK;
}

}  // namespace

int pp_assemble(const eafp_desc* d, const eafp_pp_fields* f, const eafp_pp_config* cfg, double* const out[8],
                double* residual, double* flux_ref, int on_device) {
    if (!d || !f || !cfg || !out) return EAFP_ERR_ARG;
    EAFP_TRY_RC(validate_desc(d));
    for (int q = 0; q < 8; ++q)
        if (!out[q]) return EAFP_ERR_ARG;
    const double* f3[NF3] = {f->vol,    f->Ar,    f->Ath,    f->Az,    f->alpha_l, f->ur_l,    f->uth_l,   f->uz_l,
                             f->aPr_l,  f->aPth_l, f->aPz_l, f->rho_l, f->T_l,     f->alpha_g, f->ur_g,    f->uth_g,
                             f->uz_g,   f->aPr_g,  f->aPth_g, f->aPz_g, f->rho_g,  f->T_old_g, f->p};
    for (const double* q : f3)
        if (!q) return EAFP_ERR_ARG;
    if (!f->r || !f->theta || !f->z || !f->cell_type || !f->outlet) return EAFP_ERR_ARG;
    if (!(cfg->dt > 0.0)) return EAFP_ERR_ARG;
    std::lock_guard<std::mutex> lock(g_pp_mutex);
    EAFP_TRY_RC(ensure_pp(d));
    PPWork& w = *g_pp;
    const int nr = d->nr, nth = d->ntheta, nz = d->nz;
    const long sy = nr + 4, sz = (long)(nr + 4) * (nth + 4), n = sz * (nz + 4);
    const long n2 = (long)(nr + 4) * (nth + 4);
    PPArgs a{};
    a.nr = nr;
    a.nth = nth;
    a.nz = nz;
    a.sy = sy;
    a.sz = sz;
    a.n = n;
    const double* dev3[NF3];
    if (on_device) {
        a.r = f->r;
        a.th = f->theta;
        a.z = f->z;
        for (int q = 0; q < NF3; ++q) dev3[q] = f3[q];
        a.ct = f->cell_type;
        a.outlet = f->outlet;
        a.liq_cont = f->liq_cont;
        a.pv_active = f->pv_active;
        a.gas_reach = f->gas_reach;
        a.liq_room = f->liq_room;
    } else {
        const hipMemcpyKind H2D = hipMemcpyHostToDevice;
        double* c = w.coord.p;
        EAFP_TRY(hipMemcpyAsync(c, f->r, sizeof(double) * (nr + 4), H2D, w.stream));
        EAFP_TRY(hipMemcpyAsync(c + nr + 4, f->theta, sizeof(double) * (nth + 4), H2D, w.stream));
        EAFP_TRY(hipMemcpyAsync(c + nr + nth + 8, f->z, sizeof(double) * (nz + 4), H2D, w.stream));
        a.r = c;
        a.th = c + nr + 4;
        a.z = c + nr + nth + 8;
        for (int q = 0; q < NF3; ++q) {
            EAFP_TRY(hipMemcpyAsync(w.f3.p + (long)q * n, f3[q], sizeof(double) * n, H2D, w.stream));
            dev3[q] = w.f3.p + (long)q * n;
        }
        EAFP_TRY(hipMemcpyAsync(w.ct.p, f->cell_type, sizeof(int) * n, H2D, w.stream));
        EAFP_TRY(hipMemcpyAsync(w.outlet.p, f->outlet, sizeof(int) * n2, H2D, w.stream));
        a.ct = w.ct.p;
        a.outlet = w.outlet.p;
        const int* m[3] = {f->liq_cont, f->pv_active, f->gas_reach};
        const int** dst[3] = {&a.liq_cont, &a.pv_active, &a.gas_reach};
        for (int q = 0; q < 3; ++q) {
            if (m[q]) {
                EAFP_TRY(hipMemcpyAsync(w.masks.p + (long)q * n, m[q], sizeof(int) * n, H2D, w.stream));
                *dst[q] = w.masks.p + (long)q * n;
            } else {
                *dst[q] = nullptr;
            }
        }
        if (f->liq_room) {
            EAFP_TRY(hipMemcpyAsync(w.room.p, f->liq_room, sizeof(double) * n, H2D, w.stream));
            a.liq_room = w.room.p;
        } else {
            a.liq_room = nullptr;
        }
    }
    a.vol = dev3[0]; a.Ar = dev3[1]; a.Ath = dev3[2]; a.Az = dev3[3];
    a.al = dev3[4]; a.url = dev3[5]; a.uthl = dev3[6]; a.uzl = dev3[7];
    a.aPrl = dev3[8]; a.aPthl = dev3[9]; a.aPzl = dev3[10]; a.rhol = dev3[11]; a.Tl = dev3[12];
    a.ag = dev3[13]; a.urg = dev3[14]; a.uthg = dev3[15]; a.uzg = dev3[16];
    a.aPrg = dev3[17]; a.aPthg = dev3[18]; a.aPzg = dev3[19]; a.rhog = dev3[20]; a.Toldg = dev3[21];
    a.p = dev3[22];
    a.dt = cfg->dt;
    a.rho_gas = cfg->rho_gas;
    a.T_amb = cfg->T_ambient;
    a.beta = cfg->beta_expansion;
    a.two_phase = cfg->gas_in_poisson && cfg->solve_multiphase;
    a.compress = cfg->gas_compressibility && cfg->solve_multiphase;
    a.gr = w.grads.p;
    a.gth = w.grads.p + n;
    a.gz = w.grads.p + 2 * n;
    a.gzg = w.grads.p + 3 * n;
    double** dst_out[8] = {&a.aW, &a.aE, &a.aS, &a.aN, &a.aB, &a.aT, &a.aP, &a.Su};
    for (int q = 0; q < 8; ++q) *dst_out[q] = on_device ? out[q] : w.outs.p + (long)q * n;
    a.part = w.part.p;
    const int blocks = (int)((n + BLOCK - 1) / BLOCK);
    k_pp_grad<<<blocks, BLOCK, 0, w.stream>>>(a);
    k_pp_assemble<<<blocks, BLOCK, 0, w.stream>>>(a);
    EAFP_TRY(hipGetLastError());
    if (!on_device)
        for (int q = 0; q < 8; ++q)
            EAFP_TRY(hipMemcpyAsync(out[q], w.outs.p + (long)q * n, sizeof(double) * n, hipMemcpyDeviceToHost, w.stream));
    EAFP_TRY(hipMemcpyAsync(w.part_h.data(), w.part.p, sizeof(double) * 2 * blocks, hipMemcpyDeviceToHost, w.stream));
    EAFP_TRY(hipStreamSynchronize(w.stream));
    double fl = 0.0, su = 0.0;
    for (int b = 0; b < blocks; ++b) {
        fl += w.part_h[2 * (size_t)b];
        su += w.part_h[2 * (size_t)b + 1];
    }
    if (flux_ref) *flux_ref = fl;
    if (residual) *residual = su / std::max(fl, SMALL);
    return EAFP_OK;
}

void pp_release() {
    std::lock_guard<std::mutex> lock(g_pp_mutex);
    if (g_pp) {
        if (g_pp->stream) (void)hipStreamSynchronize(g_pp->stream);
        if (g_pp->stream) (void)hipStreamDestroy(g_pp->stream);
        delete g_pp;
        g_pp = nullptr;
    }
}

}  // namespace eafp
