// tdma_line.hpp -- the pieces of the GPU line TDMA (tdma.hip) that work on
// one line, one element or one cell, shared by the kernels and by the host
// replay of the schedule in tests/tdma_check.cpp, so both run the same
// arithmetic on the same addresses.
//
// Layouts.  Each pass P (0: r lines, 1: theta lines, 2: z lines) sees the
// mesh as a grid of lines (u, v), u in 1..nu, v in 1..nv, each of n elements
// e = 0..n-1:
//   P = 0: u = j, v = k, e = i-1     (n = nr)
//   P = 1: u = i, v = k, e = j-1     (n = nth)
//   P = 2: u = i, v = j, e = k-1     (n = nz)
// The reference's loop order (v outer, u inner) makes line (u,v) wait for
// (u-1,v) and (u,v-1); the lines of one anti-diagonal w = (u-1)+(v-1) are
// independent.  Lines are numbered diagonal by diagonal,
//   l(u,v) = dstart[w] + (v - v_lo(w)),
// and every per-element array of the pass is stored element-major,
// X[e*NL + l], so the lines of one diagonal are contiguous for each e and a
// work-group reads them coalesced.  phi has rows of NLX = NL + NH: after the
// NL lines come the pass's halo lines (constant wall values, and the
// snapshot of the lines the theta wrap reads through the halo copy).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

#include "eafp_common.hpp"

namespace eafp {
namespace tdma {

// Device code addresses every array here through global-memory pointers
// (address space 1), so loads and stores are global instructions that do not
// share the LDS wait counter with the staging arrays; host code (the replay)
// sees plain pointers.  The struct layout is the same on both sides.
#if defined(__HIP_DEVICE_COMPILE__)
#define TD_GLOBAL __attribute__((address_space(1)))
#else
#define TD_GLOBAL
#endif
typedef TD_GLOBAL double* gdptr;
typedef const TD_GLOBAL double* cgdptr;
typedef const TD_GLOBAL int* cgiptr;

// Coefficient slots of the caller's arrays.
enum { CW = 0, CE, CS, CN, CB, CT, CP_, CSU, TD_NCOEF };
// Packed per-element arrays of a pass.
enum { PK_SU = 0, PK_O1, PK_O2, PK_O3, PK_O4, PK_A, PK_M, PK_C, PK_N };
// Packed per-line arrays of a pass.
enum { LN_AMF = 0, LN_APL, LN_LO, LN_HI, LN_N };

constexpr int TD_LDS_ELEMS = 5120;   // elements staged per group: d, a, m, c', 160 KiB
constexpr int TD_RING = 8;           // staged values read this many elements ahead of the recurrences

struct TdPass {
    int nu, nv, n;             // line grid and line length
    int NL, NH, NLX;           // lines, halo lines, phi row length
    cgiptr dstart;             // first line of each diagonal: nu+nv-1 entries
    gdptr pk;                  // PK_N * n * NL
    gdptr ln;                  // LN_N * NL
    gdptr phi;                 // n * NLX
};

struct TdSys {
    int nr, nth, nz, pad_;
    long sy, sz, n;            // padded strides and length
    long ncells;               // nr * nth * nz
    cgdptr c[TD_NCOEF];        // aW aE aS aN aB aT aP Su, padded
    gdptr home;                // the caller's phi (padded)
    TdPass ps[3];
};

__host__ __device__ inline long tat(const TdSys& s, int i, int j, int k) {
    return (long)(i + 1) + s.sy * (j + 1) + s.sz * (k + 1);
}

__host__ __device__ inline double clamp_pivot(double m) { return fabs(m) < SMALL ? copysign(SMALL, m) : m; }

// num / den, correctly rounded (the value of the IEEE division), in two
// parts.  On the GPU it is the hardware-assisted division written out:
// td_recip(den) is den's reciprocal with its two Newton steps, which depends
// on den only and so is computed ahead of the recurrence; td_div_fast(num,
// den, r, ok) is the numerator's part (scale, multiply, residual, fused
// correction, fix-up), the only part left on the recurrence's critical path.
// Together they are the compiler's own expansion of `/` whenever div_scale
// leaves den unchanged; td_div_fast clears ok when it does not, and the
// caller then divides with `/`.  The host (the replay) divides directly.
__host__ __device__ inline double td_recip(double den) {
#if defined(__HIP_DEVICE_COMPILE__)
    const double r0 = __builtin_amdgcn_rcp(den);
    const double r1 = fma(r0, fma(-den, r0, 1.0), r0);
    double r2 = fma(r1, fma(-den, r1, 1.0), r1);
    asm volatile("" : "+v"(r2));    // keep it here, ahead of the recurrence, not sunk into it
    return r2;
#else
    return den;
#endif
}

__host__ __device__ inline double td_div_fast(double num, double den, double r, bool& ok) {
#if defined(__HIP_DEVICE_COMPILE__)
    bool den_flag, num_flag;
    const double den_s = __builtin_amdgcn_div_scale(num, den, false, &den_flag);
    const double num_s = __builtin_amdgcn_div_scale(num, den, true, &num_flag);
    const double q = num_s * r;
    const double fm = __builtin_amdgcn_div_fmas(fma(-den, q, num_s), r, q, num_flag);
    ok = ok && __builtin_bit_cast(unsigned long long, den_s) == __builtin_bit_cast(unsigned long long, den);
    return __builtin_amdgcn_div_fixup(fm, den, num);
#else
    (void)r;
    (void)ok;
    return num / den;
#endif
}

__host__ __device__ inline void diagonal_range(int w, int nu, int nv, int& v_lo, int& v_hi) {
    v_lo = w + 2 - nu > 1 ? w + 2 - nu : 1;
    v_hi = w + 1 < nv ? w + 1 : nv;
}

__host__ __device__ inline int line_index(const TdPass& p, int u, int v) {
    const int w = u + v - 2;
    int lo, hi;
    diagonal_range(w, p.nu, p.nv, lo, hi);
    return p.dstart[w] + (v - lo);
}

// Line (u,v) of line index l (binary search over the diagonal starts).
__host__ __device__ inline void line_at(const TdPass& p, int l, int& u, int& v) {
    int lo = 0, hi = p.nu + p.nv - 2;       // the diagonal w with dstart[w] <= l < dstart[w+1]
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (p.dstart[mid] <= l) lo = mid;
        else hi = mid - 1;
    }
    int v_lo, v_hi;
    diagonal_range(lo, p.nu, p.nv, v_lo, v_hi);
    v = v_lo + (l - p.dstart[lo]);
    u = lo + 2 - v;
}

// Cell (i,j,k) of element e of line (u,v) of pass P, and back.
__host__ __device__ inline void cell_of(int P, int u, int v, int e, int& i, int& j, int& k) {
    if (P == 0) { i = e + 1; j = u; k = v; }
    else if (P == 1) { i = u; j = e + 1; k = v; }
    else { i = u; j = v; k = e + 1; }
}

__host__ __device__ inline void line_of_cell(int P, int i, int j, int k, int& u, int& v, int& e) {
    if (P == 0) { u = j; v = k; e = i - 1; }
    else if (P == 1) { u = i; v = k; e = j - 1; }
    else { u = i; v = j; e = k - 1; }
}

// Coefficient slots: the line's minus / plus neighbours and the off-line
// neighbours o1..o4 in the order the reference adds them to d.
__host__ __device__ inline void pass_slots(int P, int& am, int& ap, int* o) {
    if (P == 0) { am = CW; ap = CE; o[0] = CS; o[1] = CN; o[2] = CB; o[3] = CT; }
    else if (P == 1) { am = CS; ap = CN; o[0] = CW; o[1] = CE; o[2] = CB; o[3] = CT; }
    else { am = CB; ap = CT; o[0] = CS; o[1] = CN; o[2] = CW; o[3] = CE; }
}

// The phi-row index of the off-line neighbours o1..o4 of line (u,v).
//   P = 0 (u = j, v = k): S = (u-1,v), wrapping to (nu,v) (not yet updated
//     when line (1,v) runs); N = (u+1,v), at u = nu the snapshot of line
//     (1,v) taken at the pass start (the theta halo copy); B, T = (u,v-+1),
//     at the z walls the halo lines zlo(u), zhi(u).
//   P = 1 (u = i, v = k): W, E = (u-+1,v), at the r walls rlo(v), rhi(v);
//     B, T = (u,v-+1), at the z walls zlo(u), zhi(u).
//   P = 2 (u = i, v = j): S = (u,v-1), wrapping to (u,nv); N = (u,v+1), at
//     v = nv the snapshot of line (u,1); W, E = (u-+1,v), at the r walls
//     rlo(v), rhi(v).
__host__ __device__ inline void line_neighbours(const TdPass& p, int P, int u, int v, int* nb) {
    const int NL = p.NL, nu = p.nu, nv = p.nv;
    if (P == 0) {
        nb[0] = u > 1 ? line_index(p, u - 1, v) : line_index(p, nu, v);
        nb[1] = u < nu ? line_index(p, u + 1, v) : NL + 2 * nu + (v - 1);
        nb[2] = v > 1 ? line_index(p, u, v - 1) : NL + (u - 1);
        nb[3] = v < nv ? line_index(p, u, v + 1) : NL + nu + (u - 1);
    } else if (P == 1) {
        nb[0] = u > 1 ? line_index(p, u - 1, v) : NL + (v - 1);
        nb[1] = u < nu ? line_index(p, u + 1, v) : NL + nv + (v - 1);
        nb[2] = v > 1 ? line_index(p, u, v - 1) : NL + 2 * nv + (u - 1);
        nb[3] = v < nv ? line_index(p, u, v + 1) : NL + 2 * nv + nu + (u - 1);
    } else {
        nb[0] = v > 1 ? line_index(p, u, v - 1) : line_index(p, u, nv);
        nb[1] = v < nv ? line_index(p, u, v + 1) : NL + 2 * nv + (u - 1);
        nb[2] = u > 1 ? line_index(p, u - 1, v) : NL + (v - 1);
        nb[3] = u < nu ? line_index(p, u + 1, v) : NL + nv + (v - 1);
    }
}

// The starts and first v of the diagonals next to w, which hold every
// interior off-line neighbour of w's lines: (u-1,v) and (u,v-1) lie on w-1,
// (u+1,v) and (u,v+1) on w+1.  The same for all lines of w.
struct DiagCtx {
    int v_lo, ds_m, vlo_m, ds_p, vlo_p;
};

__host__ __device__ inline DiagCtx diag_ctx(const TdPass& p, int w) {
    DiagCtx c;
    int hi;
    diagonal_range(w, p.nu, p.nv, c.v_lo, hi);
    c.ds_m = c.vlo_m = c.ds_p = c.vlo_p = 0;
    if (w > 0) {
        c.ds_m = p.dstart[w - 1];
        diagonal_range(w - 1, p.nu, p.nv, c.vlo_m, hi);
    }
    if (w < p.nu + p.nv - 2) {
        c.ds_p = p.dstart[w + 1];
        diagonal_range(w + 1, p.nu, p.nv, c.vlo_p, hi);
    }
    return c;
}

// line_neighbours for a line of the diagonal of c, with the interior
// neighbours' indices formed from c (no per-line table); only the wrap of
// the one line at u = 1 (pass 0) or v = 1 (pass 2) looks its partner up.
__host__ __device__ inline void neighbours_in(const TdPass& p, int P, const DiagCtx& c, int u, int v, int* nb) {
    const int NL = p.NL, nu = p.nu, nv = p.nv;
    const int lm_u = c.ds_m + v - c.vlo_m;          // (u-1, v)
    const int lm_v = c.ds_m + (v - 1) - c.vlo_m;    // (u, v-1)
    const int lp_u = c.ds_p + v - c.vlo_p;          // (u+1, v)
    const int lp_v = c.ds_p + (v + 1) - c.vlo_p;    // (u, v+1)
    if (P == 0) {
        nb[0] = u > 1 ? lm_u : line_index(p, nu, v);
        nb[1] = u < nu ? lp_u : NL + 2 * nu + (v - 1);
        nb[2] = v > 1 ? lm_v : NL + (u - 1);
        nb[3] = v < nv ? lp_v : NL + nu + (u - 1);
    } else if (P == 1) {
        nb[0] = u > 1 ? lm_u : NL + (v - 1);
        nb[1] = u < nu ? lp_u : NL + nv + (v - 1);
        nb[2] = v > 1 ? lm_v : NL + 2 * nv + (u - 1);
        nb[3] = v < nv ? lp_v : NL + 2 * nv + nu + (u - 1);
    } else {
        nb[0] = v > 1 ? lm_v : line_index(p, u, nv);
        nb[1] = v < nv ? lp_v : NL + 2 * nv + (u - 1);
        nb[2] = u > 1 ? lm_u : NL + (v - 1);
        nb[3] = u < nu ? lp_u : NL + nv + (v - 1);
    }
}

// Halo lines of pass P filled once per call from the caller's wall halos:
// the line (u,v) at a wall fills the halo line that wall's neighbours read.
__host__ __device__ inline void fill_wall_lines(const TdSys& s, int P, int u, int v) {
    const TdPass& p = s.ps[P];
    const int NL = p.NL, nu = p.nu, nv = p.nv;
    for (int e = 0; e < p.n; ++e) {
        gdptr row = p.phi + (long)e * p.NLX;
        if (P == 0) {          // zlo(j), zhi(j): phi(i, j, 0), phi(i, j, nz+1)
            if (v == 1) row[NL + (u - 1)] = s.home[tat(s, e + 1, u, 0)];
            if (v == nv) row[NL + nu + (u - 1)] = s.home[tat(s, e + 1, u, s.nz + 1)];
        } else if (P == 1) {   // rlo(k), rhi(k): phi(0, j, k), phi(nr+1, j, k); zlo(i), zhi(i)
            if (u == 1) row[NL + (v - 1)] = s.home[tat(s, 0, e + 1, v)];
            if (u == nu) row[NL + nv + (v - 1)] = s.home[tat(s, s.nr + 1, e + 1, v)];
            if (v == 1) row[NL + 2 * nv + (u - 1)] = s.home[tat(s, u, e + 1, 0)];
            if (v == nv) row[NL + 2 * nv + nu + (u - 1)] = s.home[tat(s, u, e + 1, s.nz + 1)];
        } else {               // rlo(j), rhi(j): phi(0, j, k), phi(nr+1, j, k)
            if (u == 1) row[NL + (v - 1)] = s.home[tat(s, 0, v, e + 1)];
            if (u == nu) row[NL + nv + (v - 1)] = s.home[tat(s, s.nr + 1, v, e + 1)];
        }
    }
}

// Everything per call for line (u,v) of pass P: the packed right-hand-side
// coefficients, a(e) = -a_minus(e), the pivots m(e) and c'(e) of the
// reference's forward elimination (they do not depend on phi), the end
// coefficients and wall values, the wall halo lines, and for pass 0 the
// line's initial phi.
__host__ __device__ inline void setup_line(const TdSys& s, int P, int u, int v) {
    const TdPass& p = s.ps[P];
    const int l = line_index(p, u, v);
    const long NL = p.NL, n = p.n, plane = n * NL;
    int am, ap, o[4];
    pass_slots(P, am, ap, o);
    double cp = 0.0;
    for (int e = 0; e < p.n; ++e) {
        int i, j, k;
        cell_of(P, u, v, e, i, j, k);
        const long c = tat(s, i, j, k);
        gdptr x = p.pk + (long)e * NL + l;
        x[PK_SU * plane] = s.c[CSU][c];
#pragma unroll
        for (int q = 0; q < 4; ++q) x[(PK_O1 + q) * plane] = s.c[o[q]][c];
        const double a = -s.c[am][c];
        double m;
        if (e == 0) m = clamp_pivot(s.c[CP_][c]);
        else m = clamp_pivot(s.c[CP_][c] - a * cp);
        cp = -s.c[ap][c] / m;
        x[PK_A * plane] = a;
        x[PK_M * plane] = m;
        x[PK_C * plane] = cp;
        if (P == 0) p.phi[(long)e * p.NLX + l] = s.home[c];
    }
    int i0, j0, k0, i1, j1, k1;
    cell_of(P, u, v, 0, i0, j0, k0);
    cell_of(P, u, v, p.n - 1, i1, j1, k1);
    p.ln[LN_AMF * NL + l] = s.c[am][tat(s, i0, j0, k0)];
    p.ln[LN_APL * NL + l] = s.c[ap][tat(s, i1, j1, k1)];
    // wall values beyond the two ends (pass 1 reads its own line instead)
    if (P == 0) {
        p.ln[LN_LO * NL + l] = s.home[tat(s, 0, u, v)];
        p.ln[LN_HI * NL + l] = s.home[tat(s, s.nr + 1, u, v)];
    } else if (P == 2) {
        p.ln[LN_LO * NL + l] = s.home[tat(s, u, v, 0)];
        p.ln[LN_HI * NL + l] = s.home[tat(s, u, v, s.nz + 1)];
    } else {
        p.ln[LN_LO * NL + l] = 0.0;
        p.ln[LN_HI * NL + l] = 0.0;
    }
    fill_wall_lines(s, P, u, v);
}

// d(e) of line l of pass P: Su + o1 phi1 + o2 phi2, then + o3 phi3 + o4 phi4
// (the reference's operand order in all three passes), then the end terms
// d(1) += a_minus(1) phi(before), d(n) += a_plus(n) phi(after).  Pass 1 reads
// the values beyond its ends from its own line (the theta halo copy of the
// other end, not yet updated).
template <int P>
__host__ __device__ inline double rhs_elem(const TdPass& p, int l, int e, const int* nb) {
    const long NL = p.NL, plane = (long)p.n * NL, x = (long)e * NL + l;
    cgdptr row = p.phi + (long)e * p.NLX;
    double d = p.pk[PK_SU * plane + x] + p.pk[PK_O1 * plane + x] * row[nb[0]] + p.pk[PK_O2 * plane + x] * row[nb[1]];
    d = d + p.pk[PK_O3 * plane + x] * row[nb[2]] + p.pk[PK_O4 * plane + x] * row[nb[3]];
    if (e == 0) {
        const double lo = P == 1 ? p.phi[(long)(p.n - 1) * p.NLX + l] : p.ln[LN_LO * NL + l];
        d = d + p.ln[LN_AMF * NL + l] * lo;
    }
    if (e == p.n - 1) {
        const double hi = P == 1 ? p.phi[l] : p.ln[LN_HI * NL + l];
        d = d + p.ln[LN_APL * NL + l] * hi;
    }
    return d;
}

// The staged values of element (l, e): d, a, m, c' into slot `at` of the
// four staging arrays (LDS on the GPU).
template <int P>
__host__ __device__ inline void stage_elem(const TdPass& p, int l, int e, const int* nb, double* sd, double* sa,
                                           double* sm, double* sc, int at) {
    const long plane = (long)p.n * p.NL, x = (long)e * p.NL + l;
    sd[at] = rhs_elem<P>(p, l, e, nb);
    sa[at] = p.pk[PK_A * plane + x];
    sm[at] = p.pk[PK_M * plane + x];
    sc[at] = p.pk[PK_C * plane + x];
}

// The Thomas recurrences of line l from its staged values (element e at
// slot e*stride + q): d'(0) = d(0)/m(0), d'(e) = (d(e) - a(e) d'(e-1))/m(e)
// into the d slots; x(n-1) = d'(n-1), x(e) = d'(e) - c'(e) x(e+1) into the
// d slots too (the work-group then copies them to phi, coalesced).  Each chunk of TD_RING elements is read from the staging arrays,
// and its reciprocals formed, before its recurrences run, so the reads and
// the denominator-only work overlap the chain of divisions.
__host__ __device__ inline void solve_staged(const TdPass& p, int l, double* sd, const double* sa,
                                             const double* sm, const double* sc, int q, int stride) {
    const int n = p.n;
    double dp = sd[q] / sm[q];
    sd[q] = dp;
    int e0 = 1;
    for (; e0 + TD_RING <= n; e0 += TD_RING) {     // full chunks: no per-element bounds
        double cd[TD_RING], ca[TD_RING], cm[TD_RING], cr[TD_RING], out[TD_RING];
#pragma unroll
        for (int r = 0; r < TD_RING; ++r) {
            const int at = (e0 + r) * stride + q;
            cd[r] = sd[at];
            ca[r] = sa[at];
            cm[r] = sm[at];
        }
#pragma unroll
        for (int r = 0; r < TD_RING; ++r) cr[r] = td_recip(cm[r]);
        const double dp_in = dp;
        bool ok = true;
#pragma unroll
        for (int r = 0; r < TD_RING; ++r) {
            dp = td_div_fast(cd[r] - ca[r] * dp, cm[r], cr[r], ok);
            out[r] = dp;
        }
        if (!ok) {      // a rescaled division in this chunk: redo it with `/`
            dp = dp_in;
#pragma unroll
            for (int r = 0; r < TD_RING; ++r) {
                dp = (cd[r] - ca[r] * dp) / cm[r];
                out[r] = dp;
            }
        }
#pragma unroll
        for (int r = 0; r < TD_RING; ++r) sd[(e0 + r) * stride + q] = out[r];
    }
    for (int e = e0; e < n; ++e) {
        const int at = e * stride + q;
        dp = (sd[at] - sa[at] * dp) / sm[at];
        sd[at] = dp;
    }
    double x = dp;
    for (int e1 = n - 2; e1 >= 0; e1 -= TD_RING) {
        double bd[TD_RING], bc[TD_RING];
#pragma unroll
        for (int r = 0; r < TD_RING; ++r) {
            const int e = e1 - r;
            bd[r] = e >= 0 ? sd[e * stride + q] : 0.0;
            bc[r] = e >= 0 ? sc[e * stride + q] : 0.0;
        }
#pragma unroll
        for (int r = 0; r < TD_RING; ++r) {
            const int e = e1 - r;
            if (e >= 0) {
                x = bd[r] - bc[r] * x;
                sd[e * stride + q] = x;
            }
        }
    }
}

// Element `at` = e*gm + q of a staged group of gm lines from l0 on: its x,
// left in the d slots by solve_staged, copied to phi.
__host__ __device__ inline void store_staged(const TdPass& p, int l0, int gm, const double* sd, int at) {
    const int q = at % gm, e = at / gm;
    p.phi[(long)e * p.NLX + l0 + q] = sd[at];
}

// Lines of diagonal w taken per group so that group lines x n elements fit
// the staging arrays.
__host__ __device__ inline int group_lines(int n) {
    const int g = TD_LDS_ELEMS / n;
    return g > 0 ? g : 1;
}

// The snapshot halo lines of a pass, taken at its start (the theta halo
// copy): pass 0 copies line (1,v) to NL + 2nu + v-1, pass 2 line (u,1) to
// NL + 2nv + u-1.  Index t runs over n x (nv or nu) values.
__host__ __device__ inline void snapshot_value(const TdPass& p, int P, long t) {
    if (P == 0) {
        const int v = (int)(t % p.nv) + 1, e = (int)(t / p.nv);
        gdptr row = p.phi + (long)e * p.NLX;
        row[p.NL + 2 * p.nu + (v - 1)] = row[line_index(p, 1, v)];
    } else {
        const int u = (int)(t % p.nu) + 1, e = (int)(t / p.nu);
        gdptr row = p.phi + (long)e * p.NLX;
        row[p.NL + 2 * p.nv + (u - 1)] = row[line_index(p, u, 1)];
    }
}

// Periodic source plane of theta plane j (the serial copy's order: with
// nth = 1 every halo plane holds plane 1).
__host__ __device__ inline int theta_src(int j, int nth) {
    int q = (j - 1) % nth;
    if (q < 0) q += nth;
    return q + 1;
}

// Cell c (interior, r fastest) moved from pass P's phi to pass Q's; Q = 3
// is the caller's padded array.  Moving pass 1's values to pass 2 also
// writes the caller's theta halo planes of the cell's column (the copy the
// reference makes before the z lines, the last copy of a call).
__host__ __device__ inline void move_cell(const TdSys& s, int P, int Q, long c) {
    const int i = (int)(c % s.nr) + 1;
    const long rest = c / s.nr;
    const int j = (int)(rest % s.nth) + 1, k = (int)(rest / s.nth) + 1;
    int u, v, e;
    line_of_cell(P, i, j, k, u, v, e);
    const TdPass& a = s.ps[P];
    const double val = a.phi[(long)e * a.NLX + line_index(a, u, v)];
    if (Q == 3) {
        s.home[tat(s, i, j, k)] = val;
        return;
    }
    const TdPass& b = s.ps[Q];
    line_of_cell(Q, i, j, k, u, v, e);
    b.phi[(long)e * b.NLX + line_index(b, u, v)] = val;
    if (P == 1 && Q == 2) {
        const int halo[4] = {0, -1, s.nth + 1, s.nth + 2};
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (theta_src(halo[q], s.nth) == j) s.home[tat(s, i, halo[q], k)] = val;
    }
}

// The theta halo planes of the caller's array on the r/z wall columns
// (i or k in a halo): copies of the constant wall values, in the serial
// copy's order.  Index t runs over (nr+4) x (nz+4) columns.
__host__ __device__ inline void wall_halo_column(const TdSys& s, long t) {
    const int i = (int)(t % (s.nr + 4)) - 1, k = (int)(t / (s.nr + 4)) - 1;
    if (i >= 1 && i <= s.nr && k >= 1 && k <= s.nz) return;
    const int halo[4] = {0, -1, s.nth + 1, s.nth + 2};
    const int src[4] = {s.nth, s.nth - 1, 1, 2};
#pragma unroll
    for (int q = 0; q < 4; ++q) s.home[tat(s, i, halo[q], k)] = s.home[tat(s, i, src[q], k)];
}

}  // namespace tdma
}  // namespace eafp
