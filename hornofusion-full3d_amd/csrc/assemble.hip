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
    int chunk;               // blocks per XCD run (cell_of_thread)
    double *Fs[3], *ac[3];   // liquid +face exports (r, theta, z); Fs[0] == nullptr: none
    int* lk[3];
};

template <class A>
__device__ __forceinline__ bool act_l(const A& a, long o) {
    return a.liq_cont ? a.liq_cont[o] != 0 : a.al[o] >= ALPHA_FLOW_CUTOFF;
}

template <class A>
__device__ __forceinline__ bool act_g(const A& a, long o) {
    bool g = a.ag[o] >= ALPHA_FLOW_CUTOFF;
    if (a.pv_active) g = g && a.pv_active[o] != 0;
    if (a.gas_reach) g = g || (a.gas_reach[o] != 0 && a.ag[o] < ALPHA_FLOW_CUTOFF && a.ct[o] != 0);
    return g;
}

template <class A>
__device__ __forceinline__ bool pv_cell(const A& a, long o) { return a.ct[o] != 0 && (act_l(a, o) || act_g(a, o)); }

// Cell o's fluid flag and both activity flags, the same logic as above with
// every load issued up front (no short-circuit between loads), so the loads
// of several cells overlap instead of forming a chain.
struct CellAct {
    bool fluid, l, g;
};

template <class A>
__device__ __forceinline__ CellAct cell_act(const A& a, long o) {
    const int ct = a.ct[o];
    const double al = a.al[o], ag = a.ag[o];
    const int lc = a.liq_cont ? a.liq_cont[o] : 0;
    const int pv = a.pv_active ? a.pv_active[o] : 1;
    const int gr = a.gas_reach ? a.gas_reach[o] : 0;
    CellAct c;
    c.fluid = ct != 0;
    c.l = a.liq_cont ? lc != 0 : al >= ALPHA_FLOW_CUTOFF;
    bool g = ag >= ALPHA_FLOW_CUTOFF;
    if (a.pv_active) g = g && pv != 0;
    if (a.gas_reach) g = g || (gr != 0 && ag < ALPHA_FLOW_CUTOFF && ct != 0);
    c.g = g;
    return c;
}

__device__ __forceinline__ bool pv_of(const CellAct& c) { return c.fluid && (c.l || c.g); }

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

// Padded position of cell o (the host keeps every padded length below 2^31,
// so 32-bit division suffices).
template <class A>
__device__ __forceinline__ Pos padded_pos(const A& a, long o) {
    const unsigned u = (unsigned)o, sz = (unsigned)a.sz, sy = (unsigned)a.sy;
    const unsigned k = u / sz;
    const unsigned rem = u - k * sz;
    const unsigned j = rem / sy;
    const unsigned i = rem - j * sy;
    return {(int)i - 1, (int)j - 1, (int)k - 1};
}

// Padded cell of this thread.  The grid is a multiple of 8*chunk blocks,
// chunk = the blocks of two z-planes.  Block b runs on XCD x = b mod 8 and
// the XCD's t-th block (t = b/8) takes logical block
// ((t / chunk)*8 + x)*chunk + t mod chunk: each XCD sweeps runs of two
// planes (the plane below the current one is in its own L2) and the runs
// are dealt round-robin, so the XCDs share the heavy (fluid) and light
// (solid, halo) planes evenly.
template <class A>
__device__ __forceinline__ long cell_of_thread(const A& a) {
    const unsigned b = blockIdx.x, ch = (unsigned)a.chunk;
    const unsigned t = b >> 3, sc = t / ch;
    return (long)((sc * 8u + (b & 7u)) * ch + (t - sc * ch)) * BLOCK + threadIdx.x;
}

inline int pp_chunk(long sz) { return (int)(2 * ((sz + BLOCK - 1) / BLOCK)); }

// Blocks for n padded cells of plane size sz: a multiple of 8*chunk.
inline int pp_blocks(long n, long sz) {
    const long q = 8L * pp_chunk(sz);
    return (int)(((n + BLOCK - 1) / BLOCK + q - 1) / q * q);
}

template <class A>
__device__ __forceinline__ long at(const A& a, int i, int j, int k) {
    return (long)(i + 1) + a.sy * (j + 1) + a.sz * (k + 1);
}

__global__ __launch_bounds__(BLOCK) void k_pp_grad(PPArgs a) {
    const long o = cell_of_thread(a);
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
        const CellAct cW = cell_act(a, oW), cE = cell_act(a, oE), cS = cell_act(a, oS);
        const CellAct cN = cell_act(a, oN), cB = cell_act(a, oB), cT = cell_act(a, oT);
        gr = pgrad(p[oW], p[o], p[oE], a.r[i], a.r[i + 1], a.r[i + 2], i > 1 && pv_of(cW), i < a.nr && pv_of(cE));
        gth = pgrad(p[oS], p[o], p[oN], ri * a.th[j], ri * a.th[j + 1], ri * a.th[j + 2], pv_of(cS), pv_of(cN));
        const bool okm = k > 1 && pv_of(cB), okp = k < a.nz && pv_of(cT);
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
    double Fs[3], ac[3];     // the liquid's +face exports of this cell
    int lk[3];
};

// The six faces of cell o for one phase (src/mod_pressure_3d.f90:486-607).
// nb: the activity of the neighbours W, E, S, N, B, T.
__device__ void add_phase(const PPArgs& a, const Phase& ph, int i, int j, int k, long o, const CellAct* nb, Acc& c) {
    const bool gas = ph.gas;
    bool lk[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) lk[q] = nb[q].fluid && (gas ? nb[q].g : nb[q].l);
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
    if (lk[0]) {
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
    if (lk[1]) {
        d_f = 0.5 * (aef[o] * a.vol[o] / fmax(ph.aPr[o], SMALL) + aef[oE] * a.vol[oE] / fmax(ph.aPr[oE], SMALL));
        af = 0.5 * (aef[o] + aef[oE]);
        delta = a.r[i + 2] - a.r[i + 1];
        u_f = 0.5 * (ph.ur[o] + ph.ur[oE]) + d_f * (0.5 * (gr[o] + gr[oE]) - (pf[oE] - pf[o]) / delta);
        if (face_open(a, gas, u_f, o, oE)) {
            c.aE = c.aE + af * d_f * a.Ar[o] / delta;
            c.Su = c.Su - af * u_f * a.Ar[o];
            c.flux += fabs(af * u_f * a.Ar[o]);
            if (!gas) {
                c.Fs[0] = af * u_f * a.Ar[o];
                c.ac[0] = af * d_f * a.Ar[o] / delta;
                c.lk[0] = 1;
            }
        }
    }
    if (lk[2]) {
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
    if (lk[3]) {
        d_f = 0.5 * (aef[o] * a.vol[o] / fmax(ph.aPth[o], SMALL) + aef[oN] * a.vol[oN] / fmax(ph.aPth[oN], SMALL));
        af = 0.5 * (aef[o] + aef[oN]);
        delta = a.r[i + 1] * (a.th[j + 2] - a.th[j + 1]);
        u_f = 0.5 * (ph.uth[o] + ph.uth[oN]) + d_f * (0.5 * (gth[o] + gth[gN]) - (pf[oN] - pf[o]) / delta);
        if (face_open(a, gas, u_f, o, oN)) {
            c.aN = c.aN + af * d_f * a.Ath[o] / delta;
            c.Su = c.Su - af * u_f * a.Ath[o];
            c.flux += fabs(af * u_f * a.Ath[o]);
            if (!gas) {
                c.Fs[1] = af * u_f * a.Ath[o];
                c.ac[1] = af * d_f * a.Ath[o] / delta;
                c.lk[1] = 1;
            }
        }
    }
    if (lk[4]) {
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
    if (lk[5]) {
        d_f = 0.5 * (aef[o] * a.vol[o] / fmax(ph.aPz[o], SMALL) + aef[oT] * a.vol[oT] / fmax(ph.aPz[oT], SMALL));
        af = 0.5 * (aef[o] + aef[oT]);
        delta = a.z[k + 2] - a.z[k + 1];
        u_f = 0.5 * (ph.uz[o] + ph.uz[oT]) +
              d_f * (0.5 * (gz[o] + gz[oT]) - (pf[oT] - pf[o]) / delta - (gas ? liq_weight(a, o, oT) : 0.0));
        if (face_open(a, gas, u_f, o, oT)) {
            c.aT = c.aT + af * d_f * a.Az[o] / delta;
            c.Su = c.Su - af * u_f * a.Az[o];
            c.flux += fabs(af * u_f * a.Az[o]);
            if (!gas) {
                c.Fs[2] = af * u_f * a.Az[o];
                c.ac[2] = af * d_f * a.Az[o] / delta;
                c.lk[2] = 1;
            }
        }
    }
}

__global__ __launch_bounds__(BLOCK) void k_pp_assemble(PPArgs a) {
    __shared__ double sh[2 * (BLOCK / 64)];
    const long o = cell_of_thread(a);
    double acc[2] = {0.0, 0.0};
    if (o < a.n) {
        const Pos q = padded_pos(a, o);
        const int i = q.i, j = q.j, k = q.k;
        const bool interior = i >= 1 && i <= a.nr && j >= 1 && j <= a.nth && k >= 1 && k <= a.nz;
        Acc c = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0, 0, 0}};
        double aP = 0.0;
        if (interior && a.ct[o] != 0) {
            const bool al = act_l(a, o), ag = act_g(a, o);
            const CellAct nb[6] = {cell_act(a, at(a, i - 1, j, k)), cell_act(a, at(a, i + 1, j, k)),
                                   cell_act(a, at(a, i, j - 1, k)), cell_act(a, at(a, i, j + 1, k)),
                                   cell_act(a, at(a, i, j, k - 1)), cell_act(a, at(a, i, j, k + 1))};
            if (al) {
                const Phase L = {a.al, a.url, a.uthl, a.uzl, a.aPrl, a.aPthl, a.aPzl, a.gz, false};
                add_phase(a, L, i, j, k, o, nb, c);
            }
            if (a.two_phase && ag) {
                const Phase Gp = {a.ag, a.urg, a.uthg, a.uzg, a.aPrg, a.aPthg, a.aPzg, a.gzg, true};
                add_phase(a, Gp, i, j, k, o, nb, c);
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
        if (a.Fs[0]) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                a.Fs[d][o] = c.Fs[d];
                a.ac[d][o] = c.ac[d];
                a.lk[d][o] = c.lk[d];
            }
        }
    }
    block_sum<2>(acc, sh);
    if (threadIdx.x == 0) {
        a.part[2 * (long)blockIdx.x] = acc[0];
        a.part[2 * (long)blockIdx.x + 1] = acc[1];
    }
}

// ---------------------------------------------------------------------------
// Post-solve consumers of p' (src/mod_pressure_3d.f90:324-397 and
// correct_velocities :689-746), one thread per padded cell:
//   k_pp_correct  the conservative liquid fluxes of the E/N/T faces and
//                 their link flags (theta halo cells take the value of their
//                 periodic source, which is what the serial halo copy leaves
//                 there), the velocity corrections of both phases in place,
//                 and p + alpha_p pp clamped to +-P_HYDRO_CAP into a scratch
//                 array
//   k_pp_seal     the final pressure: sealed cells (fluid, no phase active)
//                 take the mean of their neighbours with an active phase,
//                 theta halo cells the value of their periodic source
// Neither kernel reduces anything, so every output is the double a
// sequential evaluation gives.  Bytes per padded cell: fluxes 6 double + 3
// int reads, 3 double + 3 int writes; velocities 2 x (3 aP + 3 u) reads and
// 3 u writes; pressure 2 reads, 1 write, then 1 read + 1 write.
constexpr double P_HYDRO_CAP = 2.0e6;          // src/mod_pressure_3d.f90:74

struct PCArgs {
    int nr, nth, nz;
    long sy, sz, n;
    const double *r, *th, *z, *vol;
    const int* ct;
    const double *al, *ag;
    const double* aP[6];       // aPr, aPth, aPz of the liquid, then of the gas
    const int *liq_cont, *pv_active, *gas_reach;
    const double* pp;
    const double *Fs[3], *ac[3];
    const int* lk_in[3];
    double* Fc[3];
    int* lk[3];
    double* vel[6];            // ur, uth, uz of the liquid, then of the gas (in place)
    const double* p;           // the pressure before the update
    double* p_upd;             // p + alpha_p pp, clamped (scratch)
    double* p_out;             // the final pressure (may be p itself)
    double alpha_p;
    int two_phase;
    int chunk;                 // blocks per XCD run (cell_of_thread)
};

// Periodic source plane of theta plane j (j itself for 1 <= j <= nth).  With
// nth = 1 the serial copy's order leaves plane 1 in every halo plane.
__device__ __forceinline__ int theta_src(int j, int nth) {
    int s = (j - 1) % nth;
    if (s < 0) s += nth;
    return s + 1;
}

__global__ __launch_bounds__(BLOCK) void k_pp_correct(PCArgs a) {
    const long o = cell_of_thread(a);
    if (o >= a.n) return;
    const Pos q = padded_pos(a, o);
    const int i = q.i, j = q.j, k = q.k;
    const bool rz = i >= 1 && i <= a.nr && k >= 1 && k <= a.nz;
    const int js = theta_src(j, a.nth);
    const long s = at(a, i, js, k);
    const bool fluid = rz && a.ct[s] != 0;
    // conservative fluxes (:333-351)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const int lk = a.lk_in[d][s];
        double fc = 0.0;
        if (fluid && lk == 1) {
            const long nb = d == 0 ? at(a, i + 1, js, k) : d == 1 ? at(a, i, js + 1, k) : at(a, i, js, k + 1);
            fc = a.Fs[d][s] + a.ac[d][s] * (a.pp[s] - a.pp[nb]);
        }
        a.Fc[d][o] = fc;
        a.lk[d][o] = lk;
    }
    // correct_velocities (:689-746): the liquid, then the gas
    if (fluid && s == o) {
        const long oW = at(a, i - 1, j, k), oE = at(a, i + 1, j, k);
        const long oS = at(a, i, j - 1, k), oN = at(a, i, j + 1, k);
        const long oB = at(a, i, j, k - 1), oT = at(a, i, j, k + 1);
        const CellAct co = cell_act(a, o);
        for (int ph = 0; ph < (a.two_phase ? 2 : 1); ++ph) {
            if (!(ph ? co.g : co.l)) continue;
            const double* alpha = ph ? a.ag : a.al;
            const double *aPr = a.aP[3 * ph], *aPth = a.aP[3 * ph + 1], *aPz = a.aP[3 * ph + 2];
            double *ur = a.vel[3 * ph], *uth = a.vel[3 * ph + 1], *uz = a.vel[3 * ph + 2];
            if (!(i == 1 || i == a.nr) && fabs(aPr[o]) > SMALL) {
                const double dc = alpha[o] * a.vol[o] / aPr[o];
                ur[o] = ur[o] - dc * (a.pp[oE] - a.pp[oW]) / (a.r[i + 2] - a.r[i]);
            }
            if (fabs(aPth[o]) > SMALL) {
                const double dc = alpha[o] * a.vol[o] / aPth[o];
                uth[o] = uth[o] - dc * (a.pp[oN] - a.pp[oS]) / (a.r[i + 1] * (a.th[j + 2] - a.th[j]));
            }
            if (!(k == 1 || k == a.nz) && fabs(aPz[o]) > SMALL) {
                const double dc = alpha[o] * a.vol[o] / aPz[o];
                uz[o] = uz[o] - dc * (a.pp[oT] - a.pp[oB]) / (a.z[k + 2] - a.z[k]);
            }
        }
    }
    // pressure update over the whole array (:358-360)
    a.p_upd[o] = fmax(-P_HYDRO_CAP, fmin(P_HYDRO_CAP, a.p[o] + a.alpha_p * a.pp[o]));
}

__global__ __launch_bounds__(BLOCK) void k_pp_seal(PCArgs a) {
    const long o = cell_of_thread(a);
    if (o >= a.n) return;
    const Pos q = padded_pos(a, o);
    const int i = q.i, k = q.k;
    const int js = theta_src(q.j, a.nth);
    const long s = at(a, i, js, k);
    double v = a.p_upd[s];
    const bool rz = i >= 1 && i <= a.nr && k >= 1 && k <= a.nz;
    if (rz) {
        const CellAct cs = cell_act(a, s);
        if (cs.fluid && !cs.l && !cs.g) {
            // sealed cell (:373-397): the neighbours W, E, S, N, B, T in this order
            const long nb[6] = {at(a, i - 1, js, k), at(a, i + 1, js, k), at(a, i, js - 1, k),
                                at(a, i, js + 1, k), at(a, i, js, k - 1), at(a, i, js, k + 1)};
            CellAct cb[6];
            double pb[6];
#pragma unroll
            for (int m = 0; m < 6; ++m) {
                cb[m] = cell_act(a, nb[m]);
                pb[m] = a.p_upd[nb[m]];
            }
            double psum = 0.0;
            int nnb = 0;
#pragma unroll
            for (int m = 0; m < 6; ++m) {
                if (!pv_of(cb[m])) continue;
                psum = psum + pb[m];
                nnb = nnb + 1;
            }
            v = nnb > 0 ? psum / (double)nnb : 0.0;
        }
    }
    a.p_out[o] = v;
}

struct PPWork {
    int nr = -1, nth = -1, nz = -1;
    hipStream_t stream = nullptr;
    DevBuf<double> coord, f3, grads, outs, part, room, exd;
    DevBuf<int> ct, outlet, masks, lki;
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
    if (n > (1L << 31) - 8L * BLOCK) return EAFP_ERR_UNSUPPORTED;   // 32-bit cell positions
    w.nr = w.nth = w.nz = -1;   // a failed allocation below leaves the workspace invalid, not stale
    EAFP_TRY_RC(w.coord.alloc((size_t)(d->nr + d->ntheta + d->nz + 12)));
    EAFP_TRY_RC(w.f3.alloc((size_t)n * NF3));
    EAFP_TRY_RC(w.grads.alloc((size_t)n * 4));
    EAFP_TRY_RC(w.outs.alloc((size_t)n * 8));
    EAFP_TRY_RC(w.room.alloc((size_t)n));
    EAFP_TRY_RC(w.ct.alloc((size_t)n));
    EAFP_TRY_RC(w.masks.alloc((size_t)n * 3));
    EAFP_TRY_RC(w.outlet.alloc((size_t)(d->nr + 4) * (d->ntheta + 4)));
    EAFP_TRY_RC(w.exd.alloc((size_t)n * 6));
    EAFP_TRY_RC(w.lki.alloc((size_t)n * 6));
    const long blocks = pp_blocks(n, (long)(d->nr + 4) * (d->ntheta + 4));
    EAFP_TRY_RC(w.part.alloc((size_t)blocks * 2));
    w.part_h.assign((size_t)blocks * 2, 0.0);
    w.nr = d->nr;
    w.nth = d->ntheta;
    w.nz = d->nz;
    return EAFP_OK;
}

}  // namespace

namespace {

// Argument checks of eafp_pp_assemble (before any device call); fills the
// export pointers (all NULL without exports).
int pp_assemble_check(const eafp_desc* d, const eafp_pp_fields* f, const eafp_pp_config* cfg, double* const out[8],
                      const eafp_pp_exports* ex, double* ex_d[6], int* ex_i[3]) {
    if (!d || !f || !cfg || !out) return EAFP_ERR_ARG;
    EAFP_TRY_RC(validate_desc(d));
    for (int q = 0; q < 8; ++q)
        if (!out[q]) return EAFP_ERR_ARG;
    for (int q = 0; q < 6; ++q) ex_d[q] = nullptr;
    for (int q = 0; q < 3; ++q) ex_i[q] = nullptr;
    if (ex) {
        double* const ed[6] = {ex->Fs_r, ex->Fs_th, ex->Fs_z, ex->ac_r, ex->ac_th, ex->ac_z};
        int* const ei[3] = {ex->lk_r, ex->lk_th, ex->lk_z};
        for (int q = 0; q < 6; ++q)
            if (!(ex_d[q] = ed[q])) return EAFP_ERR_ARG;
        for (int q = 0; q < 3; ++q)
            if (!(ex_i[q] = ei[q])) return EAFP_ERR_ARG;
    }
    const double* f3[NF3] = {f->vol,    f->Ar,    f->Ath,    f->Az,    f->alpha_l, f->ur_l,    f->uth_l,   f->uz_l,
                             f->aPr_l,  f->aPth_l, f->aPz_l, f->rho_l, f->T_l,     f->alpha_g, f->ur_g,    f->uth_g,
                             f->uz_g,   f->aPr_g,  f->aPth_g, f->aPz_g, f->rho_g,  f->T_old_g, f->p};
    for (const double* q : f3)
        if (!q) return EAFP_ERR_ARG;
    if (!f->r || !f->theta || !f->z || !f->cell_type || !f->outlet) return EAFP_ERR_ARG;
    if (!(cfg->dt > 0.0)) return EAFP_ERR_ARG;
    return EAFP_OK;
}

// The kernels' arguments apart from inputs: the mesh and the config values.
void pp_args_common(PPArgs& a, const eafp_desc* d, const eafp_pp_config* cfg) {
    a.nr = d->nr;
    a.nth = d->ntheta;
    a.nz = d->nz;
    a.sy = d->nr + 4;
    a.sz = (long)(d->nr + 4) * (d->ntheta + 4);
    a.n = a.sz * (d->nz + 4);
    a.dt = cfg->dt;
    a.rho_gas = cfg->rho_gas;
    a.T_amb = cfg->T_ambient;
    a.beta = cfg->beta_expansion;
    a.two_phase = cfg->gas_in_poisson && cfg->solve_multiphase;
    a.compress = cfg->gas_compressibility && cfg->solve_multiphase;
    a.chunk = pp_chunk(a.sz);
}

// The input pointers, from an array of the 23 three-dimensional fields.
void pp_args_fields(PPArgs& a, const double* const dev3[NF3]) {
    a.vol = dev3[0]; a.Ar = dev3[1]; a.Ath = dev3[2]; a.Az = dev3[3];
    a.al = dev3[4]; a.url = dev3[5]; a.uthl = dev3[6]; a.uzl = dev3[7];
    a.aPrl = dev3[8]; a.aPthl = dev3[9]; a.aPzl = dev3[10]; a.rhol = dev3[11]; a.Tl = dev3[12];
    a.ag = dev3[13]; a.urg = dev3[14]; a.uthg = dev3[15]; a.uzg = dev3[16];
    a.aPrg = dev3[17]; a.aPthg = dev3[18]; a.aPzg = dev3[19]; a.rhog = dev3[20]; a.Toldg = dev3[21];
    a.p = dev3[22];
}

// Every input and output in device memory, where the caller holds it.
void pp_args_device(PPArgs& a, const eafp_pp_fields* f, double* const out[8], double* const ex_d[6],
                    int* const ex_i[3]) {
    const double* f3[NF3] = {f->vol,    f->Ar,    f->Ath,    f->Az,    f->alpha_l, f->ur_l,    f->uth_l,   f->uz_l,
                             f->aPr_l,  f->aPth_l, f->aPz_l, f->rho_l, f->T_l,     f->alpha_g, f->ur_g,    f->uth_g,
                             f->uz_g,   f->aPr_g,  f->aPth_g, f->aPz_g, f->rho_g,  f->T_old_g, f->p};
    pp_args_fields(a, f3);
    a.r = f->r;
    a.th = f->theta;
    a.z = f->z;
    a.ct = f->cell_type;
    a.outlet = f->outlet;
    a.liq_cont = f->liq_cont;
    a.pv_active = f->pv_active;
    a.gas_reach = f->gas_reach;
    a.liq_room = f->liq_room;
    double** dst_out[8] = {&a.aW, &a.aE, &a.aS, &a.aN, &a.aB, &a.aT, &a.aP, &a.Su};
    for (int q = 0; q < 8; ++q) *dst_out[q] = out[q];
    for (int dd = 0; dd < 3; ++dd) {
        a.Fs[dd] = ex_d[dd];
        a.ac[dd] = ex_d[3 + dd];
        a.lk[dd] = ex_i[dd];
    }
}

}  // namespace

int pp_assemble(const eafp_desc* d, const eafp_pp_fields* f, const eafp_pp_config* cfg, double* const out[8],
                double* residual, double* flux_ref, const eafp_pp_exports* ex, int on_device) {
    double* ex_d[6];
    int* ex_i[3];
    EAFP_TRY_RC(pp_assemble_check(d, f, cfg, out, ex, ex_d, ex_i));
    std::lock_guard<std::mutex> lock(g_pp_mutex);
    EAFP_TRY_RC(ensure_pp(d));
    PPWork& w = *g_pp;
    PPArgs a{};
    pp_args_common(a, d, cfg);
    const int nr = d->nr, nth = d->ntheta;
    const long n = a.n, n2 = (long)(nr + 4) * (nth + 4);
    if (on_device) {
        pp_args_device(a, f, out, ex_d, ex_i);
    } else {
        const double* f3[NF3] = {f->vol,    f->Ar,    f->Ath,    f->Az,    f->alpha_l, f->ur_l,    f->uth_l,
                                 f->uz_l,   f->aPr_l, f->aPth_l, f->aPz_l, f->rho_l,   f->T_l,     f->alpha_g,
                                 f->ur_g,   f->uth_g, f->uz_g,   f->aPr_g, f->aPth_g,  f->aPz_g,   f->rho_g,
                                 f->T_old_g, f->p};
        const hipMemcpyKind H2D = hipMemcpyHostToDevice;
        double* c = w.coord.p;
        EAFP_TRY(hipMemcpyAsync(c, f->r, sizeof(double) * (nr + 4), H2D, w.stream));
        EAFP_TRY(hipMemcpyAsync(c + nr + 4, f->theta, sizeof(double) * (nth + 4), H2D, w.stream));
        EAFP_TRY(hipMemcpyAsync(c + nr + nth + 8, f->z, sizeof(double) * (d->nz + 4), H2D, w.stream));
        a.r = c;
        a.th = c + nr + 4;
        a.z = c + nr + nth + 8;
        const double* dev3[NF3];
        for (int q = 0; q < NF3; ++q) {
            EAFP_TRY(hipMemcpyAsync(w.f3.p + (long)q * n, f3[q], sizeof(double) * n, H2D, w.stream));
            dev3[q] = w.f3.p + (long)q * n;
        }
        pp_args_fields(a, dev3);
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
        double** dst_out[8] = {&a.aW, &a.aE, &a.aS, &a.aN, &a.aB, &a.aT, &a.aP, &a.Su};
        for (int q = 0; q < 8; ++q) *dst_out[q] = w.outs.p + (long)q * n;
        if (ex)
            for (int dd = 0; dd < 3; ++dd) {
                a.Fs[dd] = w.exd.p + (long)dd * n;
                a.ac[dd] = w.exd.p + (long)(3 + dd) * n;
                a.lk[dd] = w.lki.p + (long)dd * n;
            }
    }
    a.gr = w.grads.p;
    a.gth = w.grads.p + n;
    a.gz = w.grads.p + 2 * n;
    a.gzg = w.grads.p + 3 * n;
    a.part = w.part.p;
    const int blocks = pp_blocks(n, a.sz);
    k_pp_grad<<<blocks, BLOCK, 0, w.stream>>>(a);
    k_pp_assemble<<<blocks, BLOCK, 0, w.stream>>>(a);
    EAFP_TRY(hipGetLastError());
    if (!on_device) {
        const hipMemcpyKind D2H = hipMemcpyDeviceToHost;
        for (int q = 0; q < 8; ++q)
            EAFP_TRY(hipMemcpyAsync(out[q], w.outs.p + (long)q * n, sizeof(double) * n, D2H, w.stream));
        if (ex) {
            for (int q = 0; q < 6; ++q)
                EAFP_TRY(hipMemcpyAsync(ex_d[q], w.exd.p + (long)q * n, sizeof(double) * n, D2H, w.stream));
            for (int q = 0; q < 3; ++q)
                EAFP_TRY(hipMemcpyAsync(ex_i[q], w.lki.p + (long)q * n, sizeof(int) * n, D2H, w.stream));
        }
    }
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

int pp_correct(const eafp_desc* d, const eafp_pp_fields* f, const eafp_pp_config* cfg, double alpha_p,
               const double* pp, const eafp_pp_exports* ex, const eafp_pp_update* io, int on_device) {
    if (!d || !f || !cfg || !pp || !ex || !io) return EAFP_ERR_ARG;
    EAFP_TRY_RC(validate_desc(d));
    // inputs in staging order: vol, alpha_l, alpha_g, aP of both phases
    const double* fin[9] = {f->vol, f->alpha_l, f->alpha_g, f->aPr_l, f->aPth_l, f->aPz_l, f->aPr_g, f->aPth_g, f->aPz_g};
    for (const double* q : fin)
        if (!q) return EAFP_ERR_ARG;
    if (!f->r || !f->theta || !f->z || !f->cell_type) return EAFP_ERR_ARG;
    const double* exd[6] = {ex->Fs_r, ex->Fs_th, ex->Fs_z, ex->ac_r, ex->ac_th, ex->ac_z};
    const int* exi[3] = {ex->lk_r, ex->lk_th, ex->lk_z};
    double* const fc[3] = {io->Fc_r, io->Fc_th, io->Fc_z};
    int* const lko[3] = {io->lk_r, io->lk_th, io->lk_z};
    double* const vel[6] = {io->ur_l, io->uth_l, io->uz_l, io->ur_g, io->uth_g, io->uz_g};
    for (int q = 0; q < 6; ++q)
        if (!exd[q] || !vel[q]) return EAFP_ERR_ARG;
    for (int q = 0; q < 3; ++q)
        if (!exi[q] || !fc[q] || !lko[q]) return EAFP_ERR_ARG;
    if (!io->p) return EAFP_ERR_ARG;
    std::lock_guard<std::mutex> lock(g_pp_mutex);
    EAFP_TRY_RC(ensure_pp(d));
    PPWork& w = *g_pp;
    const int nr = d->nr, nth = d->ntheta, nz = d->nz;
    PCArgs a{};
    a.nr = nr;
    a.nth = nth;
    a.nz = nz;
    a.sy = nr + 4;
    a.sz = (long)(nr + 4) * (nth + 4);
    a.n = a.sz * (nz + 4);
    const long n = a.n;
    a.alpha_p = alpha_p;
    a.two_phase = cfg->gas_in_poisson && cfg->solve_multiphase;
    a.p_upd = w.grads.p + 3 * n;
    const double* din[9];
    if (on_device) {
        a.r = f->r;
        a.th = f->theta;
        a.z = f->z;
        for (int q = 0; q < 9; ++q) din[q] = fin[q];
        a.ct = f->cell_type;
        a.liq_cont = f->liq_cont;
        a.pv_active = f->pv_active;
        a.gas_reach = f->gas_reach;
        a.pp = pp;
        for (int q = 0; q < 3; ++q) {
            a.Fs[q] = exd[q];
            a.ac[q] = exd[3 + q];
            a.lk_in[q] = exi[q];
            a.Fc[q] = fc[q];
            a.lk[q] = lko[q];
        }
        for (int q = 0; q < 6; ++q) a.vel[q] = vel[q];
        a.p = io->p;
        a.p_out = io->p;
    } else {
        const hipMemcpyKind H2D = hipMemcpyHostToDevice;
        const size_t nd = sizeof(double) * n, ni = sizeof(int) * n;
        double* c = w.coord.p;
        EAFP_TRY(hipMemcpyAsync(c, f->r, sizeof(double) * (nr + 4), H2D, w.stream));
        EAFP_TRY(hipMemcpyAsync(c + nr + 4, f->theta, sizeof(double) * (nth + 4), H2D, w.stream));
        EAFP_TRY(hipMemcpyAsync(c + nr + nth + 8, f->z, sizeof(double) * (nz + 4), H2D, w.stream));
        a.r = c;
        a.th = c + nr + 4;
        a.z = c + nr + nth + 8;
        // f3 slots: 0..8 the inputs, 9..14 the velocities, 15 p, 16 pp
        for (int q = 0; q < 9; ++q) {
            EAFP_TRY(hipMemcpyAsync(w.f3.p + (long)q * n, fin[q], nd, H2D, w.stream));
            din[q] = w.f3.p + (long)q * n;
        }
        for (int q = 0; q < 6; ++q) {
            a.vel[q] = w.f3.p + (long)(9 + q) * n;
            EAFP_TRY(hipMemcpyAsync(a.vel[q], vel[q], nd, H2D, w.stream));
        }
        double* p_d = w.f3.p + 15 * n;
        EAFP_TRY(hipMemcpyAsync(p_d, io->p, nd, H2D, w.stream));
        a.p = p_d;
        a.p_out = p_d;
        EAFP_TRY(hipMemcpyAsync(w.f3.p + 16 * n, pp, nd, H2D, w.stream));
        a.pp = w.f3.p + 16 * n;
        EAFP_TRY(hipMemcpyAsync(w.ct.p, f->cell_type, ni, H2D, w.stream));
        a.ct = w.ct.p;
        const int* m[3] = {f->liq_cont, f->pv_active, f->gas_reach};
        const int** dst[3] = {&a.liq_cont, &a.pv_active, &a.gas_reach};
        for (int q = 0; q < 3; ++q) {
            if (m[q]) {
                EAFP_TRY(hipMemcpyAsync(w.masks.p + (long)q * n, m[q], ni, H2D, w.stream));
                *dst[q] = w.masks.p + (long)q * n;
            } else {
                *dst[q] = nullptr;
            }
        }
        for (int q = 0; q < 6; ++q) EAFP_TRY(hipMemcpyAsync(w.exd.p + (long)q * n, exd[q], nd, H2D, w.stream));
        for (int q = 0; q < 3; ++q) {
            EAFP_TRY(hipMemcpyAsync(w.lki.p + (long)q * n, exi[q], ni, H2D, w.stream));
            a.Fs[q] = w.exd.p + (long)q * n;
            a.ac[q] = w.exd.p + (long)(3 + q) * n;
            a.lk_in[q] = w.lki.p + (long)q * n;
            a.Fc[q] = w.grads.p + (long)q * n;
            a.lk[q] = w.lki.p + (long)(3 + q) * n;
        }
    }
    a.vol = din[0];
    a.al = din[1];
    a.ag = din[2];
    for (int q = 0; q < 6; ++q) a.aP[q] = din[3 + q];
    const int blocks = pp_blocks(n, a.sz);
    a.chunk = pp_chunk(a.sz);
    k_pp_correct<<<blocks, BLOCK, 0, w.stream>>>(a);
    k_pp_seal<<<blocks, BLOCK, 0, w.stream>>>(a);
    EAFP_TRY(hipGetLastError());
    if (!on_device) {
        const hipMemcpyKind D2H = hipMemcpyDeviceToHost;
        const size_t nd = sizeof(double) * n, ni = sizeof(int) * n;
        for (int q = 0; q < 3; ++q) {
            EAFP_TRY(hipMemcpyAsync(fc[q], a.Fc[q], nd, D2H, w.stream));
            EAFP_TRY(hipMemcpyAsync(lko[q], a.lk[q], ni, D2H, w.stream));
        }
        for (int q = 0; q < 6; ++q) EAFP_TRY(hipMemcpyAsync(vel[q], a.vel[q], nd, D2H, w.stream));
        EAFP_TRY(hipMemcpyAsync(io->p, a.p_out, nd, D2H, w.stream));
    }
    EAFP_TRY(hipStreamSynchronize(w.stream));
    return EAFP_OK;
}

// ---------------------------------------------------------------------------
// Allocation-free assembly (eafp_pp_assemble_noalloc).  The kernels' scratch
// -- the four gradient arrays and the continuity residual's block partials
// and sums -- lives in `scratch`, a static device array of the library sized
// for EAFP_NOALLOC_MAX_CELLS padded cells (32.5 MiB; each GPU the library
// runs on holds one until the process ends).  Host code never names the
// array itself, only its device address from hipGetSymbolAddress, taken
// once per device with that device current.  Every call on a GPU shares its
// array, which keeps the last call's contents: the kernels write every
// value they later read (every padded cell's gradients, every block's
// partials), so nothing needs clearing; calls that overlap on one GPU race.
// ---------------------------------------------------------------------------
namespace {

constexpr long NA_CELLS = EAFP_NOALLOC_MAX_CELLS;
constexpr long NA_BLOCKS = 32768;                        // block partial pairs
constexpr long NA_DOUBLES = 4 * NA_CELLS + 2 * NA_BLOCKS + 32;   // gradients, partials, the two sums (padded)
static_assert(NA_DOUBLES % 32 == 0, "256-byte multiple");

static __device__ __align__(256) double scratch[NA_DOUBLES];

constexpr int NA_MAX_DEVICES = 64;
std::mutex g_na_mutex;
double* g_na_ptr[NA_MAX_DEVICES] = {};

// The current device's address of `scratch` (looked up on its first use).
int noalloc_scratch(double** out) {
    int dev = -1;
    EAFP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= NA_MAX_DEVICES) return EAFP_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lock(g_na_mutex);
    if (!g_na_ptr[dev]) {
        void* p = nullptr;
        EAFP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return EAFP_ERR_STATE;
        g_na_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_na_ptr[dev];
    return EAFP_OK;
}

// The two sums of the block partials in block order (the host's order in
// pp_assemble, so the same bits): the block stages 256 pairs in LDS, one
// thread adds them in turn.
__global__ __launch_bounds__(BLOCK) void k_pp_partsum(const double* part, int blocks, double* sums) {
    __shared__ double sh[2 * BLOCK];
    double fl = 0.0, su = 0.0;
    for (int b0 = 0; b0 < blocks; b0 += BLOCK) {
        const int b = b0 + (int)threadIdx.x;
        if (b < blocks) {
            sh[2 * threadIdx.x] = part[2 * (long)b];
            sh[2 * threadIdx.x + 1] = part[2 * (long)b + 1];
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            const int m = min(BLOCK, blocks - b0);
            for (int q = 0; q < m; ++q) {
                fl += sh[2 * q];
                su += sh[2 * q + 1];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        sums[0] = fl;
        sums[1] = su;
    }
}

}  // namespace

int pp_assemble_noalloc(const eafp_desc* d, const eafp_pp_fields* f, const eafp_pp_config* cfg, double* const out[8],
                        double* residual, double* flux_ref, const eafp_pp_exports* ex, hipStream_t stream) {
    double* ex_d[6];
    int* ex_i[3];
    EAFP_TRY_RC(pp_assemble_check(d, f, cfg, out, ex, ex_d, ex_i));
    const long n = (long)(d->nr + 4) * (d->ntheta + 4) * (d->nz + 4);
    if (n > NA_CELLS) return EAFP_ERR_SIZE;
    PPArgs a{};
    pp_args_common(a, d, cfg);
    const long blocks = pp_blocks(n, a.sz);
    if (blocks > NA_BLOCKS) return EAFP_ERR_SIZE;
    double* sc = nullptr;
    EAFP_TRY_RC(noalloc_scratch(&sc));
    pp_args_device(a, f, out, ex_d, ex_i);
    a.gr = sc;
    a.gth = sc + n;
    a.gz = sc + 2 * n;
    a.gzg = sc + 3 * n;
    a.part = sc + 4 * NA_CELLS;
    double* sums = a.part + 2 * NA_BLOCKS;
    k_pp_grad<<<(int)blocks, BLOCK, 0, stream>>>(a);
    k_pp_assemble<<<(int)blocks, BLOCK, 0, stream>>>(a);
    k_pp_partsum<<<1, BLOCK, 0, stream>>>(a.part, (int)blocks, sums);
    EAFP_TRY(hipGetLastError());
    double hs[2] = {0.0, 0.0};
    EAFP_TRY(hipMemcpyAsync(hs, sums, sizeof(hs), hipMemcpyDeviceToHost, stream));
    EAFP_TRY(hipStreamSynchronize(stream));
    if (flux_ref) *flux_ref = hs[0];
    if (residual) *residual = hs[1] / std::max(hs[0], SMALL);
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
