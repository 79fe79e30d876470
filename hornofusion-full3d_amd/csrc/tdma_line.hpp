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
#include <cstring>

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
// Per element-line record of a pass for the chain kernel, 48 B in three
// planes of 16-byte pairs (pair q of element-line x at (q * n * NL + x) * 2,
// so the lanes of a wave, consecutive lines of one element, read whole
// cache lines): the right-hand side's coefficients Su, o1, o3, and the
// forward elimination's a(e), m(e), c'(e).  Per pass, k_tdma_prep writes
// into them what changes with the pass's phi: RC_O1 / RC_O3 hold the
// product itself on lines whose first / second updated neighbour is a wall
// or a wrap line, and the slots of a(0) and c'(n-1), which take no part in
// the recurrences, hold the end terms a_minus(1) phi(before) and a_plus(n)
// phi(after).  RC_P2 and RC_P4, the products with the two neighbours that
// do not change during the pass, live beside the record in the pass's plane
// of 16-byte pairs that k_tdma_prep streams out.
enum { RC_SU = 0, RC_O1, RC_O3, RC_A, RC_M, RC_C, RC_REC, RC_P2 = RC_REC, RC_P4, RC_N };
static_assert(RC_O1 == RC_SU + 1 && RC_A == RC_O3 + 1 && RC_C == RC_M + 1 && RC_P4 == RC_P2 + 1 &&
                  RC_SU % 2 == 0 && RC_O3 % 2 == 0 && RC_M % 2 == 0 && RC_REC % 2 == 0 && RC_N % 2 == 0,
              "records are pairs of 16-byte-aligned doubles");

constexpr int TD_LDS_ELEMS = 5120;   // elements staged per group: d, a, m, c', 160 KiB
constexpr int TD_RING = 8;           // staged values read this many elements ahead of the recurrences

struct TdPass {
    int nu, nv, n;             // line grid and line length
    int NL, NH, NLX;           // lines, halo lines, phi row length
    cgiptr dstart;             // first line of each diagonal: nu+nv-1 entries
    gdptr pk;                  // PK_N * n * NL
    gdptr ln;                  // LN_N * NL
    gdptr phi;                 // n * NLX
    gdptr rec;                 // RC_REC * n * NL: the records' planes (16-byte aligned), or null: k_tdma_pass
    gdptr pp;                  // 2 * n * NL: RC_P2, RC_P4 of element-line x at 2x (16-byte aligned)
    gdptr sink;                // TDC_SINK doubles the helpers' idle stores go to
    TD_GLOBAL int* uv;         // line l's (u, v) as u | v << 16 (chain kernel), written by k_tdma_setup
    int C, K;                  // chain kernel: elements per chunk, chunks per line
    int ML, by_u;              // chain kernel: x slots (min(nu, nv)), slot = u-1 (by_u) or v-1
    int NW, IPL;               // chain kernel: chain waves, ring entries per helper lane and chunk
};

// Field f of element-line x of pass p's records.
__host__ __device__ inline gdptr rec_at(const TdPass& p, int f, long x) {
    return p.rec + ((long)(f >> 1) * p.n * p.NL + x) * 2 + (f & 1);
}

struct TdSys {
    int nr, nth, nz, pad_;
    long sy, sz, n;            // padded strides and length
    long ncells;               // nr * nth * nz
    cgdptr c[TD_NCOEF];        // aW aE aS aN aB aT aP Su, padded
    gdptr home;                // the caller's phi (padded)
    TD_GLOBAL int* off_band;   // per pass: nonzero when a pivot is outside the chain kernel's band
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

// Biased exponent of x (bits 52..62).
__host__ __device__ inline unsigned td_bexp(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return ((unsigned)__double2hiint(x) >> 20) & 0x7ffu;
#else
    unsigned long long b;
    std::memcpy(&b, &x, sizeof b);
    return (unsigned)(b >> 52) & 0x7ffu;
#endif
}

// The chain kernel's division (td_div_band) holds for a divisor m with
// 2^-300 <= |m| < 2^301 (biased exponent 723..1323; td_bexp(m) - TD_M_LO
// at most TD_M_SPAN) and a numerator num that is +0 or has 2^-400 <= |num|
// < 2^401 (biased exponent 623..1423); td_num_dev(num) - the caller's check
// - is at most TD_N_SPAN exactly for such num.
constexpr unsigned TD_M_LO = 723, TD_M_SPAN = 600;
constexpr unsigned TD_N_LO = 623, TD_N_SPAN = 800;

// 0 for num = +0, else the distance of num's biased exponent above
// TD_N_LO (wrapping to a large value below it).
__host__ __device__ inline unsigned td_num_dev(double num) {
    unsigned long long b;
#if defined(__HIP_DEVICE_COMPILE__)
    b = __builtin_bit_cast(unsigned long long, num);
#else
    std::memcpy(&b, &num, sizeof b);
#endif
    return b == 0 ? 0u : ((unsigned)(b >> 52) & 0x7ffu) - TD_N_LO;
}

// num / m from r = td_recip(m): the hardware-assisted division
// without its scaling and fix-up steps (q = num r, the residual, the fused
// correction).  For m and num in the bands above the division's scale steps
// change neither operand (their exponents are less than 768 apart, neither
// is denormal or zero-divisor, the numerator is above 2^-970, the quotient
// and 1/m are normal) and the fix-up returns its input (a finite, normal or
// +0 quotient of finite operands; for num = +0 the product q = num r already
// carries the quotient's sign, and the residual and the correction are +0).
// So the result is the IEEE quotient whenever td_num_dev(num) <= TD_N_SPAN;
// the caller divides with `/` when a chunk fails.  A numerator of -0 fails
// the check: the correction would turn the quotient -0 / m (m > 0) into +0.
// A pass runs the chain kernel only when every pivot m of the pass is in
// band (k_tdma_setup flags the others, and k_tdma_pass takes those passes).
// The host (the replay) divides directly.
__host__ __device__ inline double td_div_band(double num, double m, double r) {
#if defined(__HIP_DEVICE_COMPILE__)
    const double q = num * r;
    const double res = fma(-m, q, num);
    return fma(res, r, q);
#else
    (void)r;
    return num / m;
#endif
}

// n / m and n % m for n * m < 2^32 from M = td_magic(m): the quotient is
// the high word of n * M (M = floor((2^32 - 1) / m) + 1; m = 1 takes M = 0
// and the quotient n; so does m = 0, the empty diagonal past a pass's last,
// which divides nothing).  tests/tdma_check.cpp checks it against `/` for every
// m up to 1024 and n up to 2^20 / m.
// a * b for 0 <= a, b and a * b < 2^24 (full-rate 24-bit multiply on the GPU).
__host__ __device__ inline int td_mul24(int a, int b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __mul24(a, b);
#else
    return a * b;
#endif
}

__host__ __device__ inline unsigned td_magic(unsigned m) { return m <= 1 ? 0u : 0xffffffffu / m + 1u; }

__host__ __device__ inline void td_divmod(int n, int m, unsigned M, int& q, int& r) {
#if defined(__HIP_DEVICE_COMPILE__)
    const unsigned hi = __umulhi((unsigned)n, M);
#else
    const unsigned hi = (unsigned)(((unsigned long long)(unsigned)n * M) >> 32);
#endif
    q = M == 0u ? n : (int)hi;
    r = n - td_mul24(q, m);
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

// k_tdma_pack packs the caller's coefficients into every pass's
// element-major planes: Su, the four off-line coefficients, a(e) =
// -a_minus(e), and, until setup_line replaces them with the pivots, aP in
// the m plane and a_plus in the c' plane; for pass 0 also the initial phi.
// It works in tiles of TD_PT^3 interior cells, one work-group each.  A
// source array's values over the tile are staged in LDS with rows of TD_PT
// consecutive r-cells read coalesced, then stored to every pass's plane in
// that pass's order: for each element, the tile's lines diagonal by
// diagonal, so consecutive threads write consecutive line indices (a tile
// holds up to TD_PT lines of each diagonal, contiguous in the plane).
constexpr int TD_PT = 16;
constexpr int TD_PT3 = TD_PT * TD_PT * TD_PT;
constexpr int TD_PSRC = TD_NCOEF + 1;      // the coefficients, then phi (pass 0 only)

struct PackTile {
    int i0, j0, k0;      // first cell (0-based)
    int ni, nj, nk;      // extents, at most TD_PT
};

// Tile b of system s; false past the system's last tile.
__host__ __device__ inline bool pack_tile_of(const TdSys& s, long b, PackTile& t) {
    const int ti = (s.nr + TD_PT - 1) / TD_PT, tj = (s.nth + TD_PT - 1) / TD_PT, tk = (s.nz + TD_PT - 1) / TD_PT;
    if (b >= (long)ti * tj * tk) return false;
    const int bi = (int)(b % ti), bj = (int)((b / ti) % tj), bk = (int)(b / ((long)ti * tj));
    t.i0 = bi * TD_PT;
    t.j0 = bj * TD_PT;
    t.k0 = bk * TD_PT;
    t.ni = s.nr - t.i0 < TD_PT ? s.nr - t.i0 : TD_PT;
    t.nj = s.nth - t.j0 < TD_PT ? s.nth - t.j0 : TD_PT;
    t.nk = s.nz - t.k0 < TD_PT ? s.nz - t.k0 : TD_PT;
    return true;
}

// The tile's extents in pass P's terms: lines (u, v) and elements.
__host__ __device__ inline void pack_tile_extents(int P, const PackTile& t, int& nu, int& nv, int& ne) {
    if (P == 0) { nu = t.nj; nv = t.nk; ne = t.ni; }
    else if (P == 1) { nu = t.ni; nv = t.nk; ne = t.nj; }
    else { nu = t.ni; nv = t.nj; ne = t.nk; }
}

// Thread tid of nt: pass P's table of the tile's lines in diagonal order:
// for the r-th line (lu, lv tile-local), order[r] = lu | lv << 8 and
// lidx[r] its line index in the pass.
__host__ __device__ inline void pack_tile_order(const TdSys& s, int P, const PackTile& t, int* order, int* lidx,
                                                int tid, int nt) {
    int nu, nv, ne;
    pack_tile_extents(P, t, nu, nv, ne);
    const int u0 = P == 0 ? t.j0 : t.i0, v0 = P == 2 ? t.j0 : t.k0;
    for (int q = tid; q < nu * nv; q += nt) {
        const int lu = q % nu, lv = q / nu, d = lu + lv;
        int r = 0;
        for (int dd = 0; dd < d; ++dd) {
            const int a = dd - (nu - 1) > 0 ? dd - (nu - 1) : 0, b = dd < nv - 1 ? dd : nv - 1;
            r += b - a + 1;
        }
        r += lv - (d - (nu - 1) > 0 ? d - (nu - 1) : 0);
        order[r] = lu | lv << 8;
        lidx[r] = line_index(s.ps[P], u0 + lu + 1, v0 + lv + 1);
    }
}

// Thread tid of nt: source array `src` (a coefficient, or phi) over the tile
// into buf, cell (ci, cj, ck) at (ck * TD_PT + cj) * TD_PT + ci.
__host__ __device__ inline void pack_tile_load(const TdSys& s, const PackTile& t, int src, double* buf, int tid,
                                               int nt) {
    cgdptr a = src < TD_NCOEF ? s.c[src] : s.home;
    for (int q = tid; q < TD_PT3; q += nt) {
        const int ci = q % TD_PT, cj = (q / TD_PT) % TD_PT, ck = q / (TD_PT * TD_PT);
        if (ci < t.ni && cj < t.nj && ck < t.nk) buf[q] = a[tat(s, t.i0 + ci + 1, t.j0 + cj + 1, t.k0 + ck + 1)];
    }
}

// Where coefficient `src` goes in pass P: its plane (PK_*), negated for
// a(e) = -a_minus(e) (every coefficient has one plane in every pass).
__host__ __device__ inline int pack_plane(int P, int src, bool& neg) {
    int am, ap, o[4];
    pass_slots(P, am, ap, o);
    neg = src == am;
    if (src == CSU) return PK_SU;
    if (src == CP_) return PK_M;
    if (src == am) return PK_A;
    if (src == ap) return PK_C;
    for (int q = 0; q < 4; ++q)
        if (src == o[q]) return PK_O1 + q;
    return -1;
}

// Thread tid of nt: the staged source `src` into pass P's plane (or, for
// phi, pass 0's phi rows), the tile's lines in diagonal order per element.
__host__ __device__ inline void pack_tile_store(const TdSys& s, int P, const PackTile& t, int src, const double* buf,
                                                const int* order, const int* lidx, int tid, int nt) {
    const TdPass& p = s.ps[P];
    bool neg = false;
    const int pl = src < TD_NCOEF ? pack_plane(P, src, neg) : -1;
    if (src < TD_NCOEF ? pl < 0 : P != 0) return;
    int nu, nv, ne;
    pack_tile_extents(P, t, nu, nv, ne);
    const int R = nu * nv;
    const int e0 = P == 0 ? t.i0 : P == 1 ? t.j0 : t.k0;
    for (int q = tid; q < ne * R; q += nt) {
        const int le = q / R, r = q - le * R, o = order[r], lu = o & 255, lv = o >> 8;
        const int ci = P == 0 ? le : lu, cj = P == 0 ? lu : P == 1 ? le : lv, ck = P == 2 ? le : lv;
        const double x = buf[(ck * TD_PT + cj) * TD_PT + ci];
        const long l = lidx[r], e = e0 + le;
        if (src < TD_NCOEF) p.pk[(long)pl * p.n * p.NL + e * p.NL + l] = neg ? -x : x;
        else p.phi[e * p.NLX + l] = x;
    }
}

constexpr int TD_SB = 8;   // setup_line's elements per block of loads

// Everything else per call for line (u,v) of pass P, after the pack: the
// pivots m(e) and c'(e) of the reference's forward elimination (they do not
// depend on phi) in place of aP and a_plus, the chain kernel's records, the
// end coefficients and wall values, and the wall halo lines.  The packed
// planes keep a line's element e at e*NL + l, so the threads of the lines of
// a diagonal read and write them side by side.
__host__ __device__ inline void setup_line(const TdSys& s, int P, int u, int v) {
    const TdPass& p = s.ps[P];
    const int l = line_index(p, u, v);
    const long NL = p.NL, n = p.n, plane = n * NL;
    int am, ap, o[4];
    pass_slots(P, am, ap, o);
    double cp = 0.0, a_first = 0.0, ap_last = 0.0;
    // blocks of TD_SB elements: the block's inputs are loaded before its
    // stores (each element's m and c' replace its own aP and a_plus), so one
    // memory latency covers TD_SB steps of the recurrence; with them the
    // chain kernel's records
    for (int e0 = 0; e0 < p.n; e0 += TD_SB) {
        double A[TD_SB], M[TD_SB], Cp[TD_SB], Su[TD_SB], O1[TD_SB], O3[TD_SB];
#pragma unroll
        for (int q = 0; q < TD_SB; ++q) {
            cgdptr x = p.pk + (long)(e0 + q < p.n ? e0 + q : p.n - 1) * NL + l;
            A[q] = x[PK_A * plane];
            M[q] = x[PK_M * plane];
            Cp[q] = x[PK_C * plane];
            if (p.rec) {
                Su[q] = x[PK_SU * plane];
                O1[q] = x[PK_O1 * plane];
                O3[q] = x[PK_O3 * plane];
            }
        }
#pragma unroll
        for (int q = 0; q < TD_SB; ++q) {
            const int e = e0 + q;
            if (e >= p.n) break;
            if (e == 0) a_first = A[q];
            if (e == p.n - 1) ap_last = Cp[q];
            double m;
            if (e == 0) m = clamp_pivot(M[q]);
            else m = clamp_pivot(M[q] - A[q] * cp);
            cp = -Cp[q] / m;
            const double c = e == p.n - 1 ? 0.0 : cp;      // c'(n-1) takes no part in the recurrences
            const long xe = (long)e * NL + l;
            gdptr x = p.pk + xe;
            x[PK_M * plane] = m;
            x[PK_C * plane] = c;
            if (p.rec) {
                // the chain kernel's record; a(0) and c'(n-1) take no part in
                // the recurrences: their slots carry the end terms (k_tdma_prep)
                typedef TD_GLOBAL double2* gd2ptr;
                *reinterpret_cast<gd2ptr>(rec_at(p, RC_SU, xe)) = make_double2(Su[q], O1[q]);
                *reinterpret_cast<gd2ptr>(rec_at(p, RC_O3, xe)) = make_double2(O3[q], e == 0 ? 0.0 : A[q]);
                *reinterpret_cast<gd2ptr>(rec_at(p, RC_M, xe)) = make_double2(m, c);
                if (td_bexp(m) - TD_M_LO > TD_M_SPAN) s.off_band[P] = 1;
            }
        }
    }
    if (p.uv) p.uv[l] = u | v << 16;
    p.ln[LN_AMF * NL + l] = -a_first;        // a_minus(0): the negation is exact
    p.ln[LN_APL * NL + l] = ap_last;
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

// k_tdma_move moves phi between the passes' layouts (Q = 3: into the
// caller's padded array) a tile of pack_tile_of at a time, through LDS:
// gathered in pass P's diagonal order, stored in pass Q's (or, for the
// caller's array, in its own order).  Moving pass 1's values to pass 2 also
// writes the caller's theta halo planes of the tile's columns (the copy the
// reference makes before the z lines, the last copy of a call).
//
// Thread tid of nt: pass P's phi over the tile into buf (cell order).
__host__ __device__ inline void move_tile_gather(const TdSys& s, int P, const PackTile& t, double* buf,
                                                 const int* order, const int* lidx, int tid, int nt) {
    const TdPass& p = s.ps[P];
    int nu, nv, ne;
    pack_tile_extents(P, t, nu, nv, ne);
    const int R = nu * nv, e0 = P == 0 ? t.i0 : P == 1 ? t.j0 : t.k0;
    for (int q = tid; q < ne * R; q += nt) {
        const int le = q / R, r = q - le * R, o = order[r], lu = o & 255, lv = o >> 8;
        const int ci = P == 0 ? le : lu, cj = P == 0 ? lu : P == 1 ? le : lv, ck = P == 2 ? le : lv;
        buf[(ck * TD_PT + cj) * TD_PT + ci] = p.phi[(long)(e0 + le) * p.NLX + lidx[r]];
    }
}

// Thread tid of nt: buf into pass Q's phi (Q's order and line table) or,
// for Q = 3, the caller's array; after pass 1 also the theta halo planes.
__host__ __device__ inline void move_tile_scatter(const TdSys& s, int P, int Q, const PackTile& t, const double* buf,
                                                  const int* order, const int* lidx, int tid, int nt) {
    if (Q < 3) {
        const TdPass& p = s.ps[Q];
        int nu, nv, ne;
        pack_tile_extents(Q, t, nu, nv, ne);
        const int R = nu * nv, e0 = Q == 0 ? t.i0 : Q == 1 ? t.j0 : t.k0;
        for (int q = tid; q < ne * R; q += nt) {
            const int le = q / R, r = q - le * R, o = order[r], lu = o & 255, lv = o >> 8;
            const int ci = Q == 0 ? le : lu, cj = Q == 0 ? lu : Q == 1 ? le : lv, ck = Q == 2 ? le : lv;
            p.phi[(long)(e0 + le) * p.NLX + lidx[r]] = buf[(ck * TD_PT + cj) * TD_PT + ci];
        }
    }
    if (Q < 3 && !(P == 1 && Q == 2)) return;
    const int halo[4] = {0, -1, s.nth + 1, s.nth + 2};
    for (int q = tid; q < TD_PT3; q += nt) {
        const int ci = q % TD_PT, cj = (q / TD_PT) % TD_PT, ck = q / (TD_PT * TD_PT);
        if (ci >= t.ni || cj >= t.nj || ck >= t.nk) continue;
        const int i = t.i0 + ci + 1, j = t.j0 + cj + 1, k = t.k0 + ck + 1;
        if (Q == 3) {
            s.home[tat(s, i, j, k)] = buf[q];
            continue;
        }
#pragma unroll
        for (int h = 0; h < 4; ++h)
            if (theta_src(halo[h], s.nth) == j) s.home[tat(s, i, halo[h], k)] = buf[q];
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

// ---------------------------------------------------------------------------
// The chain kernel (k_tdma_chain): one work-group per system and pass.
//
// The lines of a diagonal are independent; diagonal w + 1 needs the x of
// diagonal w, and its first forward step needs x(0), the last value of w's
// back substitution, so a pass is one chain of dependent steps: for every
// diagonal, n forward and n backward steps.  The kernel keeps that chain
// short and feeds it from LDS:
//   - chain waves (the first NW waves, one lane per line of the diagonal)
//     run only the recurrences, from LDS: forward d'(e) = (d(e) - a(e)
//     d'(e-1)) / m(e) with the division's reciprocal formed ahead
//     (td_div_band), then back substitution x(e) = d'(e) - c'(e) x(e+1);
//   - helper waves (the rest) stream the records from HBM into LDS ahead of
//     the chain, form d(e) from the previous diagonal's x, kept in LDS, and
//     copy each finished x to the pass's phi row.  Every helper issues the
//     same loads and stores for every chunk (a lane past the chunk repeats
//     its last entry; a store with nothing to copy goes to the sink), so the
//     compiler's waits for a chunk's loads count exactly the accesses issued
//     after them; a wave with no entry of a chunk skips the LDS work.
// x lives in LDS slots: the line (u,v) uses slot u-1 (by_u) or v-1, which
// it shares with one of its two updated neighbours, (u-1,v) or (u,v-1); the
// other sits in the slot before.  A neighbour outside the line grid (a wall
// or a wrap line) reads 1.0: from the row of ones before slot 0, or from
// its line's own slot, which holds the 1.0 it starts with until the first
// line of that slot is solved; the record then holds the neighbour's
// product in place of its coefficient, and product * 1.0 is the product.
// Every LDS array keeps a line's (or a slot's) elements contiguous, rows of
// odd length so that the lanes of a wave, one row each, fall in different
// banks; the chain then reads and writes its elements at fixed offsets from
// one address per row.  Each slot holds, element by element, the previous
// diagonal's x until the helpers have read it, then d'(e), then the new x.
// The forward sweep takes a line's elements in K chunks of C; each diagonal
// is K+2 phases separated by work-group barriers:
//   G             the helpers form d of chunk 0 (it needs x(0) of the
//                 diagonal before, just finished); the chain waits
//   F_0..F_{K-1}  the chain's forward chunks, while the helpers fill the
//                 other ring slot with the next chunk
//   B             the chain's whole back substitution, with c'(e) of the
//                 diagonal's lines, which the helpers copied to LDS with
//                 each chunk
// A helper prepares forward chunk f during the phase before it, from
// registers loaded when it prepared chunk f-2 (three register sets in
// rotation), so loads have two chunks to arrive.  The d of chunk k is
// formed after the previous diagonal's B and before F_k overwrites the
// slots with d'.
constexpr int TDC_BLOCK = 768;      // threads of a chain work-group (three waves per SIMD)
constexpr int TDC_SINK = 1024;      // doubles of a pass's sink, indexed modulo TDC_SINK
constexpr int TDC_IPL = 3;          // ring entries a helper lane fills per chunk, at most
constexpr int TDC_LDS = 20480;      // doubles of LDS (160 KiB)
constexpr int TDC_G = 4;            // ring entries the chain reads per group
static_assert((TDC_SINK & (TDC_SINK - 1)) == 0, "the sink index masks with TDC_SINK - 1");

// Plans the chain kernel for pass p: chunk length C (at most cmax), chunks
// K, slots ML, slot direction, chain waves NW, ring entries per helper lane
// IPL (the fewest that cover a chunk).  False when the pass cannot
// run it (lines of one element, diagonals over 256 lines, or LDS too small):
// then it runs k_tdma_pass.
__host__ __device__ inline int odd_up(int x) { return x | 1; }

constexpr int TDC_PAD = 4 * TDC_G;  // doubles before and after the arrays: reads ahead stay in the block

__host__ __device__ inline bool chain_plan(TdPass& p, int cmax) {
    if (p.n < 2 || (long)p.n * p.NLX >= (1L << 24)) return false;   // element offsets in 24-bit products
    const int ML = p.nu < p.nv ? p.nu : p.nv;
    const int NW = (ML + 63) / 64;
    if (NW > 4) return false;
    const long HL = TDC_BLOCK - 64 * NW;
    // ring (two slots of ML rows of odd_up(C) entries, d/r and a/m), c' and
    // x slots (ML rows of odd_up(n)), the row of ones and the padding fit
    // the LDS
    long C = ((long)TDC_LDS - 2 * TDC_PAD - (2L * ML + 1) * odd_up(p.n)) / (8L * ML) - 1;
    if ((long)TDC_IPL * HL / ML < C) C = (long)TDC_IPL * HL / ML;
    if (C > p.n) C = p.n;
    if (C > cmax) C = cmax;
    if (C < 1) return false;
    // one ring entry per helper lane where that costs at most 3 more chunks
    // per line (measured at 60x120x84: z lines in 8 chunks of 11 rather than
    // 5 of up to 20 took 0.5 % off one system's sweep and 3 % off batches;
    // r lines in 9 chunks of 7 rather than 5 of 12 were 1 % slower alone)
    const long C1 = HL / ML;
    if (C1 >= 1 && C1 < C && (p.n + C1 - 1) / C1 <= (p.n + C - 1) / C + 3) C = C1;
    // whole groups of the recurrences' reads where that costs no extra chunk
    const long Cg = C >= TDC_G ? C - C % TDC_G : C;
    if ((p.n + Cg - 1) / Cg == (p.n + C - 1) / C) C = Cg;
    const long K = (p.n + C - 1) / C;
    long Cb = (p.n + K - 1) / K;                               // the same K with the shortest chunks
    if (C % TDC_G == 0 && C >= TDC_G) Cb += (TDC_G - Cb % TDC_G) % TDC_G;
    p.K = (int)K;
    p.C = (int)(Cb < C ? Cb : C);
    p.ML = ML;
    p.by_u = p.nu <= p.nv;
    p.NW = NW;
    p.IPL = (int)(((long)p.C * ML + HL - 1) / HL);
    return true;
}

__host__ __device__ inline long chain_lds_doubles(const TdPass& p) {
    return 2L * TDC_PAD + 8L * odd_up(p.C) * p.ML + (2L * p.ML + 1) * odd_up(p.n);
}

// The two ring slots b = 0, 1 of each array follow each other: slot b of dr
// starts at dr + b * ring (no pointer arrays, which a run-time index would
// put in scratch memory).
struct TdLds {
    double2* dr;        // ring: {d(e), 1/m(e)} of chunk element t of line j at j*CS + t
    double2* am;        //       {a(e), m(e)}
    double* cp;         // c'(e) of line j at j*LS + e
    double* xb;         // x slots: element e of slot s at s*LS + e; slot -1 is the row of ones
    long ring;          // entries per slot: ML * CS
    int CS, LS;         // odd row lengths: odd_up(C), odd_up(n)
};

__host__ __device__ inline TdLds chain_lds(const TdPass& p, double* base) {
    TdLds L;
    L.CS = odd_up(p.C);
    L.LS = odd_up(p.n);
    L.ring = (long)p.ML * L.CS;
    L.dr = reinterpret_cast<double2*>(base + TDC_PAD);
    L.am = L.dr + 2 * L.ring;
    L.cp = base + TDC_PAD + 8 * L.ring;
    L.xb = L.cp + (long)(p.ML + 1) * L.LS;
    return L;
}

// Thread `tid` of `nt`: every x slot and the row of ones start as 1.0
// (before the first barrier of the pass).
__host__ __device__ inline void chain_init_slots(const TdPass& p, const TdLds& L, int tid, int nt) {
    for (int q = tid; q < (p.ML + 1) * L.LS; q += nt) L.xb[q - L.LS] = 1.0;
}

// The phase (w, idx) of a pass: idx 0 is G, 1..K are F_0..F_{K-1}, K+1 is
// B.  The iterator carries the diagonal's first v, line count and first
// line, advanced without division or table lookups.
struct TdTask {
    int w, idx, kind;   // kind: 0 none (G, or past the pass), 1 forward chunk, 2 back substitution
    int k, e0, ce;      // forward chunk, its first element, its length
    int v_lo, m, l0;    // the diagonal's first v, lines, first line index
    unsigned mM;        // td_magic(m)
    int pv_lo, pm, pl0; // the same of the diagonal before (pm = 0 on the first)
    unsigned pmM;
};

__host__ __device__ inline void chain_set_kind(const TdPass& p, TdTask& t) {
    const bool none = t.w >= p.nu + p.nv - 1 || t.idx == 0, fwd = t.idx <= p.K;
    const int k = none || !fwd ? 0 : t.idx - 1;
    const int e0 = k * p.C, rest = p.n - e0;
    t.kind = none ? 0 : fwd ? 1 : 2;
    t.k = k;
    t.e0 = e0;
    t.ce = none || !fwd ? 0 : rest < p.C ? rest : p.C;
}

__host__ __device__ inline TdTask chain_first(const TdPass& p) {
    TdTask t;
    t.w = 0;
    t.idx = 0;
    t.v_lo = 1;
    t.m = 1;
    t.mM = td_magic(1);
    t.l0 = 0;
    t.pv_lo = 1;
    t.pm = 0;
    t.pl0 = 0;
    t.pmM = 0;
    chain_set_kind(p, t);
    return t;
}

__host__ __device__ inline void chain_next_diagonal(const TdPass& p, TdTask& t) {
    t.pv_lo = t.v_lo;
    t.pm = t.m;
    t.pl0 = t.l0;
    t.pmM = t.mM;
    t.l0 += t.m;
    ++t.w;
    int hi;
    diagonal_range(t.w, p.nu, p.nv, t.v_lo, hi);
    t.m = hi - t.v_lo + 1;
    t.mM = td_magic((unsigned)t.m);
}

__host__ __device__ inline void chain_advance(const TdPass& p, TdTask& t) {
    if (++t.idx == p.K + 2) {
        t.idx = 0;
        chain_next_diagonal(p, t);
    }
    chain_set_kind(p, t);
}

// The forward chunks alone, in order (the helpers' tasks); past the last
// one the iterator stays on it, so its loads repeat valid addresses.
__host__ __device__ inline TdTask chain_first_fwd(const TdPass& p) {
    TdTask t = chain_first(p);
    t.idx = 1;
    chain_set_kind(p, t);
    return t;
}

__host__ __device__ inline void chain_advance_fwd(const TdPass& p, TdTask& t) {
    if (t.idx == p.K && t.w == p.nu + p.nv - 2) return;
    if (++t.idx == p.K + 1) {
        t.idx = 1;
        chain_next_diagonal(p, t);
    }
    chain_set_kind(p, t);
}

// The phase during which the helpers prepare forward chunk t: the one
// before it.
__host__ __device__ inline long chain_prep_phase(const TdPass& p, const TdTask& t) {
    return (long)t.w * (p.K + 2) + t.idx - 1;
}

__host__ __device__ inline long chain_phases(const TdPass& p) { return (long)(p.nu + p.nv - 1) * (p.K + 2); }

__host__ __device__ inline int chain_slot(const TdPass& p, int u, int v) { return p.by_u ? u - 1 : v - 1; }

// What a helper lane loads for one ring entry: the record.
struct TdPre {
    double r[RC_N];
};

// Ring entry `it` of forward chunk t: (element e0 + it / m, line it % m).
// The loads are unconditional (entries past the chunk repeat its last one
// and go unused), so every chunk issues the same loads.
template <int P>
__host__ __device__ inline void chain_load(const TdPass& p, const TdTask& t, int it, TdPre& x) {
    const int cnt = t.ce * t.m, itc = it < cnt ? it : cnt - 1;
    int tt, j;
    td_divmod(itc, t.m, t.mM, tt, j);
    const unsigned x0 = (unsigned)(td_mul24(t.e0 + tt, p.NL) + t.l0 + j);
    const unsigned plane = (unsigned)(p.n * p.NL);
    typedef const TD_GLOBAL double2* cgd2ptr;
    const cgd2ptr rp = reinterpret_cast<cgd2ptr>(p.rec) + x0;
#pragma unroll
    for (int q = 0; q < RC_REC / 2; ++q) {
        const double2 v2 = rp[q * plane];
        x.r[2 * q] = v2.x;
        x.r[2 * q + 1] = v2.y;
    }
    const double2 pq = reinterpret_cast<cgd2ptr>(p.pp)[x0];
    x.r[RC_P2] = pq.x;
    x.r[RC_P4] = pq.y;
}

// d(e) of ring entry `it` of forward chunk t, in the reference's operand
// order (rhs_elem): Su + o1 phi1 + o2 phi2 + o3 phi3 + o4 phi4, then the end
// term (element 0 and n-1; -0.0 elsewhere, which leaves every d as it is).
// phi1 and phi3, the updated neighbours, are read from the x slots s-1 and
// s of the line's slot s (1.0 for a wall or wrap line, whose RC_O1 / RC_O3
// hold the product); phi2 and phi4 have not changed since the pass began
// (RC_P2, RC_P4).  With d, 1/m(e), a(e) and m(e) into ring slot b, and
// c'(e) into the diagonal's c' array.  Entries past the chunk do nothing.
template <int P>
__host__ __device__ inline void chain_prepare(const TdPass& p, const TdLds& L, const TdTask& t, int b, int it,
                                              const TdPre& x) {
    if (it >= t.ce * t.m) return;
    int tt, j;
    td_divmod(it, t.m, t.mM, tt, j);
    const int e = t.e0 + tt;
    const int v = t.v_lo + j, u = t.w + 2 - v;
    const bool lo1 = p.by_u != (P == 2);                  // phi1 in slot s-1, phi3 in slot s (else the reverse)
    const double* xr = L.xb + td_mul24(chain_slot(p, u, v), L.LS) - L.LS + e;
    const double n1 = xr[lo1 ? 0 : L.LS], n3 = xr[lo1 ? L.LS : 0];
    const bool first = e == 0, last = e == p.n - 1;
    double d = x.r[RC_SU] + x.r[RC_O1] * n1;
    d = d + x.r[RC_P2];
    d = d + x.r[RC_O3] * n3;
    d = d + x.r[RC_P4];
    d = d + (first ? x.r[RC_A] : last ? x.r[RC_C] : -0.0);
    const double m = x.r[RC_M];
    const int at = td_mul24(b, (int)L.ring) + td_mul24(j, L.CS) + tt;
    double* drp = reinterpret_cast<double*>(L.dr + at);
    drp[0] = d;
    drp[1] = td_recip(m);
    L.am[at] = make_double2(first ? 0.0 : x.r[RC_A], m);
    L.cp[td_mul24(j, L.LS) + e] = last ? 0.0 : x.r[RC_C];
}

// Entry `it` of chunk (e0, ce) of diagonal w's x (m lines from v_lo, first
// line l0), from its x slot to the pass's phi row (chain_copy_out; a helper
// splits it into chain_copy_read, the value and its destination, and the
// store).  Entries past the chunk store its last entry again (the same value
// to the same address), and with `real` false the store goes to the sink,
// so every copy issues one store.
__host__ __device__ inline void chain_copy_read(const TdPass& p, const TdLds& L, bool real, int w, int v_lo, int m,
                                                unsigned mM, int l0, int e0, int ce, int it, double& x, gdptr& dst) {
    const int cnt = ce * m, itc = it < cnt ? it : cnt - 1;
    int tt, j;
    td_divmod(itc, m, mM, tt, j);
    const int e = e0 + tt;
    const int v = v_lo + j, u = w + 2 - v, sl = chain_slot(p, u, v);
    x = L.xb[td_mul24(sl > 0 ? sl : 0, L.LS) + e];
    dst = real ? p.phi + (unsigned)(td_mul24(e, p.NLX) + l0 + j) : p.sink + (it & (TDC_SINK - 1));
}

__host__ __device__ inline void chain_copy_out(const TdPass& p, const TdLds& L, bool real, int w, int v_lo, int m,
                                               unsigned mM, int l0, int e0, int ce, int it) {
    double x;
    gdptr dst;
    chain_copy_read(p, L, real, w, v_lo, m, mM, l0, e0, ce, it, x, dst);
    *dst = x;
}

// Keeps the compiler from moving loads past this point (device code only).
#if defined(__HIP_DEVICE_COMPILE__)
#define TD_SCHED_FENCE() __builtin_amdgcn_sched_barrier(0)
#else
#define TD_SCHED_FENCE() ((void)0)
#endif

// The redo of a forward chunk whose numerators left the band of
// td_div_band: the same steps with `/` (device code; the host divides with
// `/` throughout).
__host__ __device__ inline void chain_forward_redo(const double2* dr, const double2* am, double* xs, int ce,
                                                   double dp_in, unsigned worst, double& dp) {
#if defined(__HIP_DEVICE_COMPILE__)
    if (worst > TD_N_SPAN) {
        dp = dp_in;
        for (int tt = 0; tt < ce; ++tt) {
            const double2 dv = dr[tt], av = am[tt];
            dp = (dv.x - av.x * dp) / av.y;
            xs[tt] = dp;
        }
    }
#else
    (void)dr, (void)am, (void)xs, (void)ce, (void)dp_in, (void)worst, (void)dp;
#endif
}

// chain_forward for a chunk of NG groups of TDC_G elements (the last one
// possibly short), NG known at compile time: the same steps in the same
// order, the ring read two groups ahead into three register sets, and no
// run-time loop or tail around them.  A one-wave probe timed it at about 80
// cycles per element plus 74 per chunk, against 78 plus 438 for
// chain_forward_loop.
template <int NG>
__host__ __device__ inline void chain_forward_fixed(const TdPass& p, const TdLds& L, const TdTask& t, int b, int j,
                                                    double& dp) {
    if (j >= t.m) return;
    const int v = t.v_lo + j, u = t.w + 2 - v, s = chain_slot(p, u, v);
    const double2* dr = L.dr + b * L.ring + (long)j * L.CS;
    const double2* am = L.am + b * L.ring + (long)j * L.CS;
    double* xs = L.xb + (long)s * L.LS + t.e0;
    const int ce = t.ce;
    if (t.e0 == 0) dp = 0.0;
    const double dp_in = dp;
    unsigned worst = 0;
    double2 D[3][TDC_G], A[3][TDC_G];
    // step g reads group g (reads past the chunk land in the row's tail,
    // the next row or the padding, and go unused) and runs group g - 2
#pragma unroll
    for (int g = 0; g < NG + 2; ++g) {
        if (g < NG) {
#pragma unroll
            for (int r = 0; r < TDC_G; ++r) {
                D[g % 3][r] = dr[g * TDC_G + r];
                A[g % 3][r] = am[g * TDC_G + r];
            }
            TD_SCHED_FENCE();
        }
        if (g >= 2) {
            const int h = g - 2;
#pragma unroll
            for (int r = 0; r < TDC_G; ++r) {
                if (h < NG - 1 || h * TDC_G + r < ce) {
                    const double num = D[h % 3][r].x - A[h % 3][r].x * dp;
                    dp = td_div_band(num, A[h % 3][r].y, D[h % 3][r].y);
                    const unsigned dev = td_num_dev(num);
                    worst = dev > worst ? dev : worst;
                    xs[h * TDC_G + r] = dp;
                }
            }
        }
    }
    chain_forward_redo(dr, am, xs, ce, dp_in, worst, dp);
}

// The chain lane of line j of forward chunk t: elements e0..e0+ce-1, d'
// into the line's x slot, dp carried across chunks.  Element 0's ring entry
// holds a = +0 and dp starts at +0, so every element takes the same step:
// d - (+0)(+0) is d itself for every d, signed zeros included.  The ring is
// read TDC_G elements at a time, two groups ahead of the recurrence, so LDS
// latency stays off it; only the chunk's last group tests its bounds.  A
// chunk whose numerators leave the band of td_div_band is redone with `/`.
// Any chunk length; chain_forward takes it for chunks over TDC_FIXED_NG
// groups.
__host__ __device__ inline void chain_forward_loop(const TdPass& p, const TdLds& L, const TdTask& t, int b, int j,
                                                   double& dp) {
    if (j >= t.m) return;
    const int v = t.v_lo + j, u = t.w + 2 - v, s = chain_slot(p, u, v);
    const double2* dr = L.dr + b * L.ring + (long)j * L.CS;
    const double2* am = L.am + b * L.ring + (long)j * L.CS;
    double* xs = L.xb + (long)s * L.LS + t.e0;
    const int ce = t.ce, ng = (ce + TDC_G - 1) / TDC_G;
    if (t.e0 == 0) dp = 0.0;
    const double dp_in = dp;
    unsigned worst = 0;
    double2 d0[TDC_G], a0[TDC_G], d1[TDC_G], a1[TDC_G], d2[TDC_G], a2[TDC_G];
    // group g holds elements g*TDC_G.. (reads past the chunk land in the
    // row's tail, the next row or the padding, and go unused)
    auto load = [&](int g, double2 (&D)[TDC_G], double2 (&A)[TDC_G]) {
#pragma unroll
        for (int r = 0; r < TDC_G; ++r) {
            D[r] = dr[g * TDC_G + r];
            A[r] = am[g * TDC_G + r];
        }
    };
    auto step = [&](int tt, const double2& D, const double2& A) {
        const double num = D.x - A.x * dp;
        dp = td_div_band(num, A.y, D.y);
        const unsigned dev = td_num_dev(num);
        worst = dev > worst ? dev : worst;
        xs[tt] = dp;
    };
    auto full = [&](int g, const double2 (&D)[TDC_G], const double2 (&A)[TDC_G]) {
#pragma unroll
        for (int r = 0; r < TDC_G; ++r) step(g * TDC_G + r, D[r], A[r]);
    };
    auto part = [&](int g, const double2 (&D)[TDC_G], const double2 (&A)[TDC_G]) {
#pragma unroll
        for (int r = 0; r < TDC_G; ++r)
            if (g * TDC_G + r < ce) step(g * TDC_G + r, D[r], A[r]);
    };
    load(0, d0, a0);
    load(1, d1, a1);
    int g = 0;
    for (; g + 3 < ng; g += 3) {
        load(g + 2, d2, a2);
        full(g, d0, a0);
        load(g + 3, d0, a0);
        full(g + 1, d1, a1);
        load(g + 4, d1, a1);
        full(g + 2, d2, a2);
    }
    if (ng - g == 1) {
        part(g, d0, a0);
    } else if (ng - g == 2) {
        full(g, d0, a0);
        part(g + 1, d1, a1);
    } else {
        load(g + 2, d2, a2);
        full(g, d0, a0);
        full(g + 1, d1, a1);
        part(g + 2, d2, a2);
    }
    chain_forward_redo(dr, am, xs, ce, dp_in, worst, dp);
}

constexpr int TDC_FIXED_NG = 6;     // chunks of up to 24 elements take chain_forward_fixed

// Forward chunk t for the chain lane of line j (chain_forward_fixed or
// chain_forward_loop by the chunk's group count).
__host__ __device__ inline void chain_forward(const TdPass& p, const TdLds& L, const TdTask& t, int b, int j,
                                              double& dp) {
    switch ((t.ce + TDC_G - 1) / TDC_G) {
    case 1: chain_forward_fixed<1>(p, L, t, b, j, dp); return;
    case 2: chain_forward_fixed<2>(p, L, t, b, j, dp); return;
    case 3: chain_forward_fixed<3>(p, L, t, b, j, dp); return;
    case 4: chain_forward_fixed<4>(p, L, t, b, j, dp); return;
    case 5: chain_forward_fixed<5>(p, L, t, b, j, dp); return;
    case 6: chain_forward_fixed<6>(p, L, t, b, j, dp); return;
    default: chain_forward_loop(p, L, t, b, j, dp); return;
    }
}
static_assert(TDC_FIXED_NG == 6, "chain_forward dispatches group counts 1 to 6");

// Back substitution of line j of diagonal t from e = n-1 down to 0: x into
// the line's x slot.  c'(n-1) is +0 and x starts at +0, so the first
// element's x = d' - (+0)(+0) is d' itself and every element takes the same
// step; d'(e) and c'(e) are read TDC_G at a time, two groups ahead (reads
// below element 0 land in the row before or the padding, and go unused).
__host__ __device__ inline void chain_backward(const TdPass& p, const TdLds& L, const TdTask& t, int j) {
    if (j >= t.m) return;
    const int v = t.v_lo + j, u = t.w + 2 - v, s = chain_slot(p, u, v);
    const int n = p.n, ng = (n + TDC_G - 1) / TDC_G;
    double* xs = L.xb + (long)s * L.LS + n - 1;          // element n-1-i at xs[-i]
    const double* cp = L.cp + (long)j * L.LS + n - 1;
    double x = 0.0;
    double p0[TDC_G], c0[TDC_G], p1[TDC_G], c1[TDC_G], p2[TDC_G], c2[TDC_G];
    auto load = [&](int g, double (&X)[TDC_G], double (&Cc)[TDC_G]) {
#pragma unroll
        for (int r = 0; r < TDC_G; ++r) {
            X[r] = xs[-(g * TDC_G + r)];
            Cc[r] = cp[-(g * TDC_G + r)];
        }
    };
    auto step = [&](int i, double X, double Cc) {
        x = X - Cc * x;
        xs[-i] = x;
    };
    auto full = [&](int g, const double (&X)[TDC_G], const double (&Cc)[TDC_G]) {
#pragma unroll
        for (int r = 0; r < TDC_G; ++r) step(g * TDC_G + r, X[r], Cc[r]);
    };
    auto part = [&](int g, const double (&X)[TDC_G], const double (&Cc)[TDC_G]) {
#pragma unroll
        for (int r = 0; r < TDC_G; ++r)
            if (g * TDC_G + r < n) step(g * TDC_G + r, X[r], Cc[r]);
    };
    load(0, p0, c0);
    load(1, p1, c1);
    int g = 0;
    for (; g + 3 < ng; g += 3) {
        load(g + 2, p2, c2);
        full(g, p0, c0);
        load(g + 3, p0, c0);
        full(g + 1, p1, c1);
        load(g + 4, p1, c1);
        full(g + 2, p2, c2);
    }
    if (ng - g == 1) {
        part(g, p0, c0);
    } else if (ng - g == 2) {
        full(g, p0, c0);
        part(g + 1, p1, c1);
    } else {
        load(g + 2, p2, c2);
        full(g, p0, c0);
        full(g + 1, p1, c1);
        part(g + 2, p2, c2);
    }
}

// The per-pass products of k_tdma_prep for element e of line l: RC_P2 and
// RC_P4 (the neighbours that keep their pass-start values: (u+1,v), (u,v+1),
// walls, and the wrap's partner (1,v) / (u,1) read before the pass changes
// it), RC_O1 / RC_O3 on lines whose updated neighbour is a wall or wrap line,
// and for e = 0 and e = n-1 the line's end-term product, in the slot of a(0)
// or c'(n-1).
template <int P>
__host__ __device__ inline void prep_elem(const TdPass& p, int l, int e) {
    const int uv = p.uv[l], u = uv & 0xffff, v = uv >> 16;
    int nb[4];
    line_neighbours(p, P, u, v, nb);
    if (P == 0 && u == p.nu) nb[1] = line_index(p, 1, v);
    if (P == 2 && v == p.nv) nb[1] = line_index(p, u, 1);
    const long NL = p.NL, plane = (long)p.n * NL, x = (long)e * NL + l;
    cgdptr row = p.phi + (long)e * p.NLX;
    typedef TD_GLOBAL double2* gd2ptr;
    reinterpret_cast<gd2ptr>(p.pp)[x] =
        make_double2(p.pk[PK_O2 * plane + x] * row[nb[1]], p.pk[PK_O4 * plane + x] * row[nb[3]]);
    if (P == 2 ? v == 1 : u == 1) *rec_at(p, RC_O1, x) = p.pk[PK_O1 * plane + x] * row[nb[0]];
    if (P == 2 ? u == 1 : v == 1) *rec_at(p, RC_O3, x) = p.pk[PK_O3 * plane + x] * row[nb[2]];
    if (e == 0) {
        const double lo = P == 1 ? p.phi[(long)(p.n - 1) * p.NLX + l] : p.ln[LN_LO * NL + l];
        *rec_at(p, RC_A, x) = p.ln[LN_AMF * NL + l] * lo;           // the end term, in the slot of a(0)
    } else if (e == p.n - 1) {
        const double hi = P == 1 ? p.phi[l] : p.ln[LN_HI * NL + l];
        *rec_at(p, RC_C, x) = p.ln[LN_APL * NL + l] * hi;           // the end term, in the slot of c'(n-1)
    }
}

}  // namespace tdma
}  // namespace eafp
