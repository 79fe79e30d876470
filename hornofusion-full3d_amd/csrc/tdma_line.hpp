// tdma_line.hpp -- per-line pieces of the line TDMA (tdma.hip), shared by the
// GPU kernels and by the host replay of the schedule in tests/tdma_check.cpp,
// so both run the same arithmetic on the same addresses.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

#include "eafp_common.hpp"

namespace eafp {
namespace tdma {

// Coefficient slots.
enum { CW = 0, CE, CS, CN, CB, CT, CP_, CSU, TD_NCOEF };

struct TdSys {
    int nr, nth, nz, pad_;
    long sy, sz, n;            // padded strides and length
    const double* c[TD_NCOEF];    // aW aE aS aN aB aT aP Su
    double* phi;
    double* piv;               // m and c' of the r, theta and z passes: 6 * n
    double* dp;                // forward-elimination scratch: n
};

__host__ __device__ inline long tat(const TdSys& s, int i, int j, int k) {
    return (long)(i + 1) + s.sy * (j + 1) + s.sz * (k + 1);
}

__host__ __device__ inline double clamp_pivot(double m) { return fabs(m) < SMALL ? copysign(SMALL, m) : m; }

// Line geometry of pass p (0 r, 1 theta, 2 z) for line coordinates (u, v):
// r: (j,k), theta: (i,k), z: (i,j).  Returns the first element, the stride and
// the length; am/ap are the slots of the line's minus and plus coefficients.
struct Line {
    long o0, st;
    int n, am, ap;
};

__host__ __device__ inline Line line_of(const TdSys& s, int p, int u, int v) {
    if (p == 0) return {tat(s, 1, u, v), 1, s.nr, CW, CE};
    if (p == 1) return {tat(s, u, 1, v), s.sy, s.nth, CS, CN};
    return {tat(s, u, v, 1), s.sz, s.nz, CB, CT};
}

// Right-hand side of element o of a line of pass P, before the end terms:
// r and z lines take Su + aS phi_S + aN phi_N, then the remaining pair
// (aB, aT for r; aW, aE for z); theta lines take Su + aW phi_W + aE phi_E,
// then aB phi_B + aT phi_T.
template <int P>
__host__ __device__ inline double line_rhs(const TdSys& s, const double* __restrict__ const* c,
                                           const double* phi, long o) {
    double d;
    if (P == 0) {
        d = c[CSU][o] + c[CS][o] * phi[o - s.sy] + c[CN][o] * phi[o + s.sy];
        d = d + c[CB][o] * phi[o - s.sz] + c[CT][o] * phi[o + s.sz];
    } else if (P == 1) {
        d = c[CSU][o];
        d = d + c[CW][o] * phi[o - 1] + c[CE][o] * phi[o + 1];
        d = d + c[CB][o] * phi[o - s.sz] + c[CT][o] * phi[o + s.sz];
    } else {
        d = c[CSU][o] + c[CS][o] * phi[o - s.sy] + c[CN][o] * phi[o + s.sy];
        d = d + c[CW][o] * phi[o - 1] + c[CE][o] * phi[o + 1];
    }
    return d;
}

// One line of pass P: forward elimination into dp, back substitution into phi.
template <int P>
__host__ __device__ inline void solve_line(const TdSys& s, const double* __restrict__ const* c, int u, int v) {
    const Line L = line_of(s, P, u, v);
    double* phi = s.phi;
    double* __restrict__ dp = s.dp;
    const double* __restrict__ M = s.piv + (long)(2 * P) * s.n;
    const double* __restrict__ C = s.piv + (long)(2 * P + 1) * s.n;
    const double* __restrict__ am = c[L.am];
    const double* __restrict__ ap = c[L.ap];
    const int n = L.n;
    const long st = L.st;
    const long o0 = L.o0, o1 = L.o0 + (long)(n - 1) * st;
    // first element, with the end terms d(1) += a_m(1) phi(before), d(n) += a_p(n) phi(after)
    double d = line_rhs<P>(s, c, phi, o0);
    d = d + am[o0] * phi[o0 - st];
    if (n == 1) d = d + ap[o1] * phi[o1 + st];
    double dpv = d / M[o0];
    dp[o0] = dpv;
    long o = o0;
    for (int e = 1; e < n; ++e) {
        o += st;
        d = line_rhs<P>(s, c, phi, o);
        if (e == n - 1) d = d + ap[o1] * phi[o1 + st];
        const double a = -am[o];
        dpv = (d - a * dpv) / M[o];
        dp[o] = dpv;
    }
    double x = dpv;
    phi[o1] = x;
    o = o1;
    for (int e = n - 2; e >= 0; --e) {
        o -= st;
        x = dp[o] - C[o] * x;
        phi[o] = x;
    }
}

// Line of pass p for the pivot pass: the index t of one line among the
// nth*nz r lines, then the nr*nz theta lines, then the nr*nth z lines.
// Returns false past the last line.
__host__ __device__ inline bool pivot_line(const TdSys& s, long t, int& p, int& u, int& v) {
    const long l_r = (long)s.nth * s.nz, l_t = (long)s.nr * s.nz, l_z = (long)s.nr * s.nth;
    if (t < l_r) {
        p = 0; u = (int)(t % s.nth) + 1; v = (int)(t / s.nth) + 1;
    } else if ((t -= l_r) < l_t) {
        p = 1; u = (int)(t % s.nr) + 1; v = (int)(t / s.nr) + 1;
    } else if ((t -= l_t) < l_z) {
        p = 2; u = (int)(t % s.nr) + 1; v = (int)(t / s.nr) + 1;
    } else {
        return false;
    }
    return true;
}

// Pivots m(i) and c'(i) of line (u,v) of pass p (the reference's forward
// elimination without the right-hand side).
__host__ __device__ inline void line_pivots(const TdSys& s, int p, int u, int v) {
    const Line L = line_of(s, p, u, v);
    const double* __restrict__ b = s.c[CP_];
    const double* __restrict__ am = p == 0 ? s.c[CW] : p == 1 ? s.c[CS] : s.c[CB];
    const double* __restrict__ ap = p == 0 ? s.c[CE] : p == 1 ? s.c[CN] : s.c[CT];
    double* __restrict__ M = s.piv + (long)(2 * p) * s.n;
    double* __restrict__ C = s.piv + (long)(2 * p + 1) * s.n;
    long o = L.o0;
    double m = clamp_pivot(b[o]);
    double cp = -ap[o] / m;
    M[o] = m;
    C[o] = cp;
    for (int e = 1; e < L.n; ++e) {
        o += L.st;
        const double a = -am[o];
        m = clamp_pivot(b[o] - a * cp);
        cp = -ap[o] / m;
        M[o] = m;
        C[o] = cp;
    }
}

// The four theta halo values of column (i,k), in the order of the serial
// periodic copy (src/mod_mpi_topology.f90:303-310): planes 0, -1, nth+1,
// nth+2 from nth, nth-1, 1, 2.
__host__ __device__ inline void theta_copy_column(const TdSys& s, int i, int k) {
    const long base = tat(s, i, 0, k);
    const long src[4] = {s.nth, s.nth - 1, 1, 2};
    const long dst[4] = {0, -1, s.nth + 1, s.nth + 2};
#pragma unroll
    for (int q = 0; q < 4; ++q) s.phi[base + s.sy * dst[q]] = s.phi[base + s.sy * src[q]];
}

// Lines (u,v) of diagonal w of an nu x nv line grid: v in [v_lo, v_hi], u = w + 2 - v.
__host__ __device__ inline void diagonal_range(int w, int nu, int nv, int& v_lo, int& v_hi) {
    v_lo = w + 2 - nu > 1 ? w + 2 - nu : 1;
    v_hi = w + 1 < nv ? w + 1 : nv;
}

}  // namespace tdma
}  // namespace eafp
