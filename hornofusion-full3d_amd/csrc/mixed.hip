// mixed.hip -- mixed-precision p' solver: FP32 Jacobi-PCG on compressed
// symmetric face weights, inside FP64 iterative refinement.
//
// Setup (one pass, k_compress): for every cell the three faces it owns
// (+r, +theta, +z) get one symmetric weight w = (a_own + a_neighbour)/2,
// rounded to the storage format (FP32, BF16, or FP16 scaled by the owner's
// diagonal aP); the diagonal excess e = aP - sum(interior links) (compliance,
// acoustic, Dirichlet and wall terms; identity rows give e = aP) is stored
// in the same format, and the Jacobi inverse 1/(sum of the six rounded
// weights + rounded e) in FP32.
//
// Inner operator, flux form:  (A_lo p)_c = sum_f w_f (p_c - p_nb) + e_c p_c.
// It is an exact weighted graph Laplacian plus a diagonal: rounding the
// weights keeps it symmetric and keeps the constant null vector of every
// pure-Neumann component exact, so (1-d)A <= A_lo <= (1+d)A with d the
// storage rounding.
//
// Refinement (FP64), every `interval` inner iterations for all systems at
// once (k_res, k_mreplace): x += d, r = Su - A*x with the caller's seven
// FP64 coefficient arrays, FP64 norms, stop test, per-component sums; the
// inner residual is replaced by fl32(r - mean) on every singular
// (pure-Neumann) component and d restarts from zero.  The inner CG keeps
// its search direction across the replacement (the reliable-update form
// of refinement: no Krylov restart, so the inner iteration count stays
// close to the FP64 CG's), or restarts it (classical refinement) on
// request.  Every dot product is accumulated in FP64.
//
// Inner iteration: two kernels as in cg64.hip, FP32 vectors.
//   k_mspmv   beta, p = z + beta*p (recomputed at neighbours), q = A_lo p, p.q
//   k_mupdate alpha, d += alpha p, r -= alpha q, z = r/D, r.r and r.z
// Algorithmic bytes per cell and inner iteration: k_mspmv reads z, p_old,
// 3 weights and e (+1 FP32 scale for FP16); writes p, q: 8 + 4w (+4) + 8 B;
// k_mupdate reads d, r, p, q, 1/D, writes d, r, z: 32 B.  FP32 weights:
// 64 B, BF16: 56 B, FP16 + scale: 60 B.
//
// One captured graph runs a whole refinement step for every system:
// interval x (k_mspmv, k_mupdate), k_res, k_rsum, k_mreplace, k_mcount.
#include <cmath>
#include <vector>

#include "batch.hpp"

namespace eafp {

// Per-system state of the mixed solver, double-buffered by kernel role:
// MxDyn changes every inner iteration (k_mspmv reads slot 1 and writes
// slot 0, k_mupdate the reverse); MxStat changes once per refinement step
// (k_res copies slot 1 to slot 0, k_mreplace writes slot 1 from slot 0).
// Every thread loads the few MxDyn fields it needs and the writer stores
// its new values straight from registers.
struct MxDyn {
    double rz, alpha, beta;
    int fresh;          // next k_mspmv starts a new Krylov sequence (beta = 0)
    int broken;         // inner breakdown: idle until the next refinement step
    int total_inner;
    int done;           // mirror of MxStat::done for the inner kernels
};

struct MxStat {
    double bnorm, b1, res_l2, res_l1;
    double mean_x[MAX_COMP];
    int n_refine, done, status, xbuf;
};

struct MxConst {
    double tol;
    double inv_cnt[MAX_COMP];
    int ncomp, max_steps, restart, pad;
};

constexpr int NRES = 4 + 2 * MAX_COMP;   // r.r, |r|, Su.Su, |Su|, sum r per comp, sum x per comp

struct MxArgs {
    const TileDesc* tiles;
    int ntiles;
    int nsys;
    const ColDesc* cols;   // column segments (column-marching stencil)
    // FP64 operand
    const double *aW, *aE, *aS, *aN, *aB, *aT, *aP, *b, *fold, *x0;
    const unsigned char* comp;
    double* X[2];        // FP64 iterate, ping-pong per refinement step
    double* r64;         // FP64 residual of the current refinement step
    double* xout;        // projected result
    // compressed operand
    void* w[3];          // +r, +theta, +z face weights (format dependent)
    float* scale;        // FP16 only: per-cell scale (the owner's aP)
    void* e;             // diagonal excess, in the weight format (FP16: scaled like the weights)
    float* dinv;         // Jacobi inverse of the rounded operator
    const LineTile* lt;  // line tiles (z-line preconditioner)
    unsigned* fac;       // z-line factors per cell: FP16 l (low half), BF16 1/d (high half)
    int line;            // 1: z-line preconditioner; the r.z partials are per line tile
    int state_first;     // 1 (batches): read the system's state before the cell loads, so the
                         // blocks of converged systems load nothing; 0 (one system): overlap them
    // inner FP32 vectors
    float *d, *r, *z, *q;
    float* p[2];
    long sink;           // a 16-byte slot past the cells of p[0], p[1] and q: stores that must be issued but not kept
    double* partR;       // ntiles * NRES (k_res)
    double* sumR;        // nsys * NRES: partR reduced per system (k_rsum)
    const SysGeom* geom; // per-system geometry and tile ranges
    double* partA;       // ntiles        (k_mspmv: p.q)
    double* partB;       // ntiles * 2, or nltiles * 2 with `line` (k_mreplace, k_mupdate: r.r, r.z)
    MxDyn* dyn[2];
    MxStat* stat[2];
    const MxConst* cst;
    int* ndone;
};

// ---------------------------------------------------------------------------
// Weight storage formats
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned short f32_to_bf16_rne(float v) {
    unsigned u = __float_as_uint(v);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

template <int F>
struct WFmt;

template <>
struct WFmt<EAFP_COEF_FP32> {
    using T = float;
    __device__ static float load(const void* w, const float*, long o) { return ((const float*)w)[o]; }
    __device__ static F4 load4(const void* w, const float*, long o) { return ld4((const float*)w + o); }
    __device__ static T enc(float v, float) { return v; }
    __device__ static float dec(T h, float) { return h; }
};

template <>
struct WFmt<EAFP_COEF_BF16> {
    using T = unsigned short;
    __device__ static float load(const void* w, const float*, long o) {
        return __uint_as_float((unsigned)((const unsigned short*)w)[o] << 16);
    }
    __device__ static F4 load4(const void* w, const float*, long o) {
        unsigned short h[4];
        __builtin_memcpy(h, __builtin_assume_aligned((const unsigned short*)w + o, 8), 8);
        F4 r;
#pragma unroll
        for (int q = 0; q < 4; ++q) r.v[q] = __uint_as_float((unsigned)h[q] << 16);
        return r;
    }
    __device__ static T enc(float v, float) { return f32_to_bf16_rne(v); }
    __device__ static float dec(T h, float) { return __uint_as_float((unsigned)h << 16); }
};

template <>
struct WFmt<EAFP_COEF_FP16> {
    using T = _Float16;
    __device__ static float load(const void* w, const float* sc, long o) {
        return (float)((const _Float16*)w)[o] * sc[o];
    }
    __device__ static F4 load4(const void* w, const float* sc, long o) {
        _Float16 h[4];
        __builtin_memcpy(h, __builtin_assume_aligned((const _Float16*)w + o, 8), 8);
        F4 r = ld4(sc + o);
#pragma unroll
        for (int q = 0; q < 4; ++q) r.v[q] *= (float)h[q];
        return r;
    }
    __device__ static T enc(float v, float s) { return (_Float16)(v / s); }
    __device__ static float dec(T h, float s) { return (float)h * s; }
};

size_t wbytes(int fmt) { return fmt == EAFP_COEF_FP32 ? 4 : 2; }

// ---------------------------------------------------------------------------
// Setup: symmetric face weights, excess, Jacobi inverse.
// ---------------------------------------------------------------------------
template <int F>
__global__ __launch_bounds__(BLOCK) void k_compress(MxArgs a, int tile) {
    using W = WFmt<F>;
    const TileWalk w = my_tiles(a.ntiles);
    for (int t = w.first; t < w.end;) {
        const TileDesc T = a.tiles[t];
        const int nrun = w.run(t, T.tile_end);
        for (int u = 0; u < nrun; ++u) {
            const int base = (t + u * w.stride - T.tile_begin) * tile;
            const int c_end = min(T.n, base + tile);
            for (int c = base + (int)threadIdx.x; c < c_end; c += BLOCK) {
                const CellIdx id = decompose(c, T);
                const long o = T.off + c;
                const Nb nb = neighbours(id, T, o);
                const bool hW = nb.hW, hE = nb.hE, hB = nb.hB, hT = nb.hT;
                const long oW = nb.W, oE = nb.E, oS = nb.S, oN = nb.N, oB = nb.B, oT = nb.T;
                const double ap = a.aP[o];
                const double wr = hE ? 0.5 * (a.aE[o] + a.aW[oE]) : 0.0;
                const double wrm = hW ? 0.5 * (a.aE[oW] + a.aW[o]) : 0.0;
                const double wt = 0.5 * (a.aN[o] + a.aS[oN]);
                const double wtm = 0.5 * (a.aN[oS] + a.aS[o]);
                const double wz = hT ? 0.5 * (a.aT[o] + a.aB[oT]) : 0.0;
                const double wzm = hB ? 0.5 * (a.aT[oB] + a.aB[o]) : 0.0;
                const double e = ap - (wr + wrm + wt + wtm + wz + wzm);
                // scale of a cell's own weights (FP16 only): its diagonal
                auto sc = [&](long q) {
                    const float v = (float)a.aP[q];
                    return v > 0.f ? v : 1.f;
                };
                const float s_o = sc(o);
                const typename W::T er = W::enc((float)wr, s_o), et = W::enc((float)wt, s_o),
                                     ez = W::enc((float)wz, s_o);
                // the rounded weights exactly as the inner operator sees them
                float D = W::dec(er, s_o) + W::dec(et, s_o) + W::dec(ez, s_o);
                D += hW ? W::dec(W::enc((float)wrm, sc(oW)), sc(oW)) : 0.f;
                D += W::dec(W::enc((float)wtm, sc(oS)), sc(oS));
                D += hB ? W::dec(W::enc((float)wzm, sc(oB)), sc(oB)) : 0.f;
                const typename W::T ee = W::enc((float)e, s_o);
                D += W::dec(ee, s_o);
                ((typename W::T*)a.w[0])[o] = er;
                ((typename W::T*)a.w[1])[o] = et;
                ((typename W::T*)a.w[2])[o] = ez;
                if (F == EAFP_COEF_FP16) a.scale[o] = s_o;
                ((typename W::T*)a.e)[o] = ee;
                a.dinv[o] = D > 0.f ? 1.0f / D : 0.0f;
                if (a.line) a.q[o] = D;   // the z-line factorization reads it (q is free until the first stencil)
            }
        }
        t += nrun * w.stride;
    }
}

// ---------------------------------------------------------------------------
// Refinement, part 1 (reads slot 1, writes slot 0): x_new = x + d,
// r = Su - A*x_new with the caller's FP64 operator (cg_3d_mpi's matvec,
// src/mod_solver_3d.f90:588-605), partials of norms and per-component sums.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_res(MxArgs a, int tile) {
    __shared__ double sh[NRES * (BLOCK / 64)];
    const TileWalk w = my_tiles(a.ntiles);
    const float* __restrict__ d = a.d;
    for (int t = w.first; t < w.end;) {
        const TileDesc T = a.tiles[t];
        const int s = T.sys;
        const int nrun = w.run(t, T.tile_end);
        const int done = a.stat[1][s].done;
        const int rnd = a.stat[1][s].n_refine;
        if (t == T.tile_begin && threadIdx.x == 0) {
            const MxStat st = a.stat[1][s];
            const MxDyn dy = a.dyn[1][s];
            store_state(&a.stat[0][s], st);
            store_state(&a.dyn[0][s], dy);
        }
        if (!done) {
            const bool first = (rnd == 0);
            const double* __restrict__ xo = first ? a.x0 : a.X[rnd & 1];
            double* __restrict__ xn = a.X[(rnd + 1) & 1];
            double acc[NRES];
#pragma unroll
            for (int c = 0; c < NRES; ++c) acc[c] = 0.0;
            const double* __restrict__ aW = a.aW;
            const double* __restrict__ aE = a.aE;
            const double* __restrict__ aS = a.aS;
            const double* __restrict__ aN = a.aN;
            const double* __restrict__ aB = a.aB;
            const double* __restrict__ aT = a.aT;
            const double* __restrict__ aP = a.aP;
            const double* __restrict__ bb = a.b;
            const double* __restrict__ fold = a.fold;
            const unsigned char* __restrict__ comp = a.comp;
            double* __restrict__ r64 = a.r64;
            for (int u = 0; u < nrun; ++u) {
                const int base = (t + u * w.stride - T.tile_begin) * tile;
                const int c_end = min(T.n, base + tile);
                for (int c = base + (int)threadIdx.x; c < c_end; c += BLOCK) {
                    const CellIdx id = decompose(c, T);
                    const long o = T.off + c;
                    const Nb nb = neighbours(id, T, o);
                    const long idx[7] = {o, nb.W, nb.E, nb.S, nb.N, nb.B, nb.T};
                    double xv[7];
#pragma unroll
                    for (int q = 0; q < 7; ++q) xv[q] = first ? xo[idx[q]] : xo[idx[q]] + (double)d[idx[q]];
                    // the FP64 operand is read once per refinement: nontemporal
                    const double cP = __builtin_nontemporal_load(aP + o), cW = __builtin_nontemporal_load(aW + o),
                                 cE = __builtin_nontemporal_load(aE + o), cS = __builtin_nontemporal_load(aS + o),
                                 cN = __builtin_nontemporal_load(aN + o), cB = __builtin_nontemporal_load(aB + o),
                                 cT = __builtin_nontemporal_load(aT + o);
                    const double su = __builtin_nontemporal_load(bb + o), fo = __builtin_nontemporal_load(fold + o);
                    const int slot = comp[o];
                    const double xc = xv[0];
                    const double ax = cP * xc - cW * (nb.hW ? xv[1] : 0.0) - cE * (nb.hE ? xv[2] : 0.0) -
                                      cS * xv[3] - cN * xv[4] - cB * (nb.hB ? xv[5] : 0.0) - cT * (nb.hT ? xv[6] : 0.0);
                    const double res = su - (ax - fo);
                    xn[o] = xc;
                    r64[o] = res;
                    acc[0] += res * res;
                    acc[1] += fabs(res);
                    acc[2] += su * su;
                    acc[3] += fabs(su);
#pragma unroll
                    for (int q = 0; q < MAX_COMP; ++q) {
                        acc[4 + q] += slot == q ? res : 0.0;
                        acc[4 + MAX_COMP + q] += slot == q ? xc : 0.0;
                    }
                }
            }
            block_sum<NRES>(acc, sh);
            store_run<NRES>(a.partR, t, nrun, w.stride, acc);
        }
        t += nrun * w.stride;
    }
}

// Refinement, between the parts: one block per system reduces the k_res
// partials of its tiles (in tile order, as every other reduction) into the
// system's NRES sums, which every block of k_mreplace then reads, instead
// of each block reducing the NRES columns of all the system's tiles again.
__global__ __launch_bounds__(BLOCK) void k_rsum(MxArgs a) {
    __shared__ double sh[NRES * (BLOCK / 64)];
    const int s = blockIdx.x;
    const SysGeom g = a.geom[s];
    double tot[NRES];
    reduce_tiles<NRES>(a.partR, g.tile_begin, g.tile_end, tot, sh);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int c = 0; c < NRES; ++c) a.sumR[(long)s * NRES + c] = tot[c];
    }
}

// ---------------------------------------------------------------------------
// Refinement, part 2 (reads slot 0, writes slot 1): true residual, stop
// tests, component means; then the inner residual is replaced by
// r = fl32(r64 - mean per component) (the compatibility projection),
// z = r/D, d = 0; the search direction is kept (or zeroed on a restart).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_mreplace(MxArgs a, int tile) {
    __shared__ double sh[NRES * (BLOCK / 64)];
    const TileWalk w = my_tiles(a.ntiles);
    for (int t = w.first; t < w.end;) {
        const TileDesc T = a.tiles[t];
        const int s = T.sys;
        const int nrun = w.run(t, T.tile_end);
        const MxStat P = a.stat[0][s];
        bool done = P.done != 0;
        bool zero_p = false;
        double mean_r[MAX_COMP] = {};
        if (!done) {
            const MxConst& K = a.cst[s];
            double tot[NRES];
#pragma unroll
            for (int c = 0; c < NRES; ++c) tot[c] = a.sumR[(long)s * NRES + c];
            const int rnd = P.n_refine;
            MxStat N = P;
            N.bnorm = rnd == 0 ? sqrt(tot[2]) : P.bnorm;
            N.b1 = rnd == 0 ? tot[3] : P.b1;
            int status = EAFP_STOP_RUNNING;
            if (N.bnorm < SMALL) {
                N.res_l2 = 0.0;
                N.res_l1 = 0.0;
                status = EAFP_STOP_ZERO_RHS;
            } else {
                N.res_l2 = sqrt(tot[0]) / N.bnorm;
                N.res_l1 = N.b1 > 0.0 ? tot[1] / N.b1 : tot[1];
                if (N.res_l2 < K.tol) status = EAFP_STOP_CONVERGED;
                else if (rnd >= K.max_steps) status = EAFP_STOP_MAX_ITER;
            }
            done = status != EAFP_STOP_RUNNING;
            const MxDyn D0 = a.dyn[0][s];
            zero_p = (rnd == 0) || K.restart || D0.broken;
#pragma unroll
            for (int q = 0; q < MAX_COMP; ++q) {
                mean_r[q] = q < K.ncomp ? tot[4 + q] * K.inv_cnt[q] : 0.0;
                N.mean_x[q] = q < K.ncomp ? tot[4 + MAX_COMP + q] * K.inv_cnt[q] : 0.0;
            }
            N.xbuf = (rnd + 1) & 1;
            MxDyn D = D0;
            if (done) {
                N.done = 1;
                N.status = status;
                D.done = 1;
            } else {
                N.n_refine = rnd + 1;
                D.fresh = zero_p ? 1 : 0;
                D.broken = 0;
            }
            if (t == T.tile_begin && threadIdx.x == 0) {
                store_state(&a.stat[1][s], N);
                store_state(&a.dyn[1][s], D);
            }
        } else if (t == T.tile_begin && threadIdx.x == 0) {
            const MxDyn D0 = a.dyn[0][s];
            store_state(&a.stat[1][s], P);
            store_state(&a.dyn[1][s], D0);
        }
        if (!done) {
            const double* __restrict__ r64 = a.r64;
            const unsigned char* __restrict__ comp = a.comp;
            const float* __restrict__ dinv = a.dinv;
            float* __restrict__ r = a.r;
            float* __restrict__ z = a.z;
            float* __restrict__ d = a.d;
            float* __restrict__ p0 = a.p[0];
            double acc[2] = {0.0, 0.0};
            for (int u = 0; u < nrun; ++u) {
                const int base = (t + u * w.stride - T.tile_begin) * tile;
                const int c_end = min(T.n, base + tile);
                for (int c = base + (int)threadIdx.x; c < c_end; c += BLOCK) {
                    const long o = T.off + c;
                    const int slot = comp[o];
                    const double rv = __builtin_nontemporal_load(r64 + o);
                    const float iv = __builtin_nontemporal_load(dinv + o);
                    double mu = 0.0;
#pragma unroll
                    for (int q = 0; q < MAX_COMP; ++q) mu = slot == q ? mean_r[q] : mu;
                    const float rr = (float)(rv - mu);
                    const float zz = iv * rr;
                    __builtin_nontemporal_store(rr, r + o);
                    z[o] = zz;
                    __builtin_nontemporal_store(0.0f, d + o);
                    if (zero_p) p0[o] = 0.0f;
                    acc[0] += (double)rr * (double)rr;
                    acc[1] += (double)rr * (double)zz;
                }
            }
            block_sum<2>(acc, sh);
            store_run<2>(a.partB, t, nrun, w.stride, acc);
        }
        t += nrun * w.stride;
    }
}

// ---------------------------------------------------------------------------
// Inner iteration, role 0 (reads slot 1, writes slot 0): beta from the r.z
// partials, p = z + beta*p, q = A_lo p (flux form), p.q.
// ---------------------------------------------------------------------------
template <int F, int CPT>
__global__ __launch_bounds__(BLOCK) void k_mspmv(MxArgs a, int par) {
    using W = WFmt<F>;
    constexpr int TILE = BLOCK * CPT;
    __shared__ double sh[2 * (BLOCK / 64)];
    const float* __restrict__ pold = a.p[par];
    float* __restrict__ pnew = a.p[par ^ 1];
    const float* __restrict__ z = a.z;
    float* __restrict__ qo = a.q;
    const TileWalk w = my_tiles(a.ntiles);
    for (int t = w.first; t < w.end;) {
        const TileDesc T = a.tiles[t];
        const int s = T.sys;
        const int nrun = w.run(t, T.tile_end);
        float zv[CPT][7], pv[CPT][7], wv[CPT][6], ev[CPT];
        long oc[CPT];
        bool ok[CPT], mW[CPT], mB[CPT];
        auto load = [&](int tile) {
            const int c0 = (tile - T.tile_begin) * TILE + (int)threadIdx.x;
#pragma unroll
            for (int m = 0; m < CPT; ++m) {
                const int c = c0 + m * BLOCK;
                ok[m] = c < T.n;
                const int cc = clamp_cell(c, T.n);
                const CellIdx id = decompose(cc, T);
                const long o = T.off + cc;
                const Nb nb = neighbours(id, T, o);
                oc[m] = o;
                mW[m] = nb.hW;
                mB[m] = nb.hB;
                const long idx[7] = {o, nb.W, nb.E, nb.S, nb.N, nb.B, nb.T};
#pragma unroll
                for (int q = 0; q < 7; ++q) {
                    zv[m][q] = z[idx[q]];
                    pv[m][q] = pold[idx[q]];
                }
                // own +faces; -faces are the neighbours' +faces.  A wall
                // +face weight is stored as 0 and its neighbour index is the
                // cell itself, so it adds 0.
                wv[m][0] = W::load(a.w[0], a.scale, o);
                wv[m][1] = W::load(a.w[0], a.scale, nb.W);
                wv[m][2] = W::load(a.w[1], a.scale, o);
                wv[m][3] = W::load(a.w[1], a.scale, nb.S);
                wv[m][4] = W::load(a.w[2], a.scale, o);
                wv[m][5] = W::load(a.w[2], a.scale, nb.B);
                ev[m] = W::load(a.e, a.scale, o);
            }
        };
        // the first batch of cell loads is in flight while the partials are reduced
        MxDyn P;
        if (a.state_first) {   // a batch: converged systems load nothing
            P = a.dyn[1][s];
            if (!(P.done || P.broken)) load(t);
        } else {
            load(t);
            P = a.dyn[1][s];
        }
        MxDyn N = P;
        bool skip = P.done || P.broken;
        if (!skip) {
            double tot[2];
            reduce_tiles<2>(a.partB, a.line ? T.lt_begin : T.tile_begin, a.line ? T.lt_end : T.tile_end, tot, sh);
            if (!(tot[1] > 0.0)) {
                skip = true;     // r.z lost positivity: wait for the next refinement step
                N.broken = 1;
            } else {
                N.beta = P.fresh ? 0.0 : tot[1] / P.rz;
                N.rz = tot[1];
                N.fresh = 0;
            }
        }
        if (t == T.tile_begin && threadIdx.x == 0) store_state(&a.dyn[0][s], N);
        if (!skip) {
            const float fb = (float)N.beta;
            double acc[1] = {0.0};
            for (int u = 0; u < nrun; ++u) {
                if (u) load(t + u * w.stride);
#pragma unroll
                for (int m = 0; m < CPT; ++m) {
                    float p[7];
#pragma unroll
                    for (int q = 0; q < 7; ++q) p[q] = zv[m][q] + fb * pv[m][q];
                    const float pc = p[0];
                    float y = ev[m] * pc;
                    y += wv[m][0] * (pc - p[2]);
                    y += (mW[m] ? wv[m][1] : 0.f) * (pc - p[1]);
                    y += wv[m][2] * (pc - p[4]);
                    y += wv[m][3] * (pc - p[3]);
                    y += wv[m][4] * (pc - p[6]);
                    y += (mB[m] ? wv[m][5] : 0.f) * (pc - p[5]);
                    if (ok[m]) {
                        pnew[oc[m]] = pc;
                        qo[oc[m]] = y;
                        acc[0] += (double)pc * (double)y;
                    }
                }
            }
            block_sum<1>(acc, sh);
            store_run<1>(a.partA, t, nrun, w.stride, acc);
        }
        t += nrun * w.stride;
    }
}

// ---------------------------------------------------------------------------
// Inner iteration, role 1 (reads slot 0, writes slot 1): alpha from the
// p.q partials, d += alpha p, r -= alpha q, z = r/D, partials r.r and r.z.
// ---------------------------------------------------------------------------
template <int CPT>
__global__ __launch_bounds__(BLOCK) void k_mupdate(MxArgs a, int par) {
    constexpr int TILE = BLOCK * CPT;
    __shared__ double sh[2 * (BLOCK / 64)];
    const float* __restrict__ pv = a.p[par ^ 1];
    float* __restrict__ d = a.d;
    float* __restrict__ r = a.r;
    float* __restrict__ zo = a.z;
    const float* __restrict__ q = a.q;
    const float* __restrict__ dinv = a.dinv;
    const TileWalk w = my_tiles(a.ntiles);
    for (int t = w.first; t < w.end;) {
        const TileDesc T = a.tiles[t];
        const int s = T.sys;
        const int nrun = w.run(t, T.tile_end);
        float dv[CPT], pvv[CPT], rv[CPT], qv[CPT], iv[CPT];
        long oc[CPT];
        bool ok[CPT];
        auto load = [&](int tile) {
            const int c0 = (tile - T.tile_begin) * TILE + (int)threadIdx.x;
#pragma unroll
            for (int m = 0; m < CPT; ++m) {
                const int c = c0 + m * BLOCK;
                ok[m] = c < T.n;
                const long o = T.off + clamp_cell(c, T.n);
                oc[m] = o;
                dv[m] = __builtin_nontemporal_load(d + o);
                pvv[m] = __builtin_nontemporal_load(pv + o);
                rv[m] = __builtin_nontemporal_load(r + o);
                qv[m] = __builtin_nontemporal_load(q + o);
                iv[m] = __builtin_nontemporal_load(dinv + o);
            }
        };
        // the first batch of cell loads is in flight while the partials are reduced
        MxDyn P;
        if (a.state_first) {   // a batch: converged systems load nothing
            P = a.dyn[0][s];
            if (!(P.done || P.broken)) load(t);
        } else {
            load(t);
            P = a.dyn[0][s];
        }
        MxDyn N = P;
        bool skip = P.done || P.broken;
        if (!skip) {
            double tot[1];
            reduce_tiles<1>(a.partA, T.tile_begin, T.tile_begin + T.pa_n, tot, sh);
            if (!(fabs(tot[0]) > SMALL)) {
                skip = true;
                N.broken = 1;
            } else {
                N.alpha = P.rz / tot[0];
                N.total_inner = P.total_inner + 1;
            }
        }
        if (t == T.tile_begin && threadIdx.x == 0) store_state(&a.dyn[1][s], N);
        if (!skip) {
            const float fa = (float)N.alpha;
            double acc[2] = {0.0, 0.0};
            for (int u = 0; u < nrun; ++u) {
                if (u) load(t + u * w.stride);
#pragma unroll
                for (int m = 0; m < CPT; ++m) {
                    const float rr = rv[m] - fa * qv[m];
                    const float zz = iv[m] * rr;
                    if (ok[m]) {
                        __builtin_nontemporal_store(dv[m] + fa * pvv[m], d + oc[m]);
                        __builtin_nontemporal_store(rr, r + oc[m]);
                        zo[oc[m]] = zz;
                        acc[0] += (double)rr * (double)rr;
                        acc[1] += (double)rr * (double)zz;
                    }
                }
            }
            block_sum<2>(acc, sh);
            store_run<2>(a.partB, t, nrun, w.stride, acc);
        }
        t += nrun * w.stride;
    }
}

// ---------------------------------------------------------------------------
// Quad forms of the two inner-iteration kernels, for batches in which every
// nz is a multiple of 4.  A thread takes Q quads -- 4 consecutive cells of
// one z-row -- per tile, so every operand of a quad arrives in one 16-byte
// access (8 bytes for 2-byte weights) instead of four, and p = z + beta*p
// is formed once per loaded value.  The arithmetic per cell is that of
// k_mspmv / k_mupdate, term for term; only the order of the partial sums
// differs.  The tile table is the same (a tile is BLOCK*4*Q cells).  The
// update streams its operands with nontemporal loads and d, r with
// nontemporal stores; z, which the next stencil reads, is stored plainly.
// ---------------------------------------------------------------------------
template <int F, int Q>
__global__ __launch_bounds__(BLOCK) void k_mspmv4(MxArgs a, int par) {
    using W = WFmt<F>;
    constexpr int TILE = BLOCK * 4 * Q;
    __shared__ double sh[2 * (BLOCK / 64)];
    const float* __restrict__ pold = a.p[par];
    float* __restrict__ pnew = a.p[par ^ 1];
    const float* __restrict__ z = a.z;
    float* __restrict__ qo = a.q;
    const TileWalk w = my_tiles(a.ntiles);
    for (int t = w.first; t < w.end;) {
        const TileDesc T = a.tiles[t];
        const int s = T.sys;
        const int nrun = w.run(t, T.tile_end);
        // per quad: z and p_old at the quad and its W/E/S/N quads, at the
        // cells below and above, and the weights of its faces
        F4 zq[Q][5], pq[Q][5], wq[Q][5], ev[Q];
        float zb[Q][2], pb[Q][2], wB[Q];
        long oc[Q];
        bool ok[Q], mW[Q], mB[Q];
        auto load = [&](int tile) {
            const int c0 = (tile - T.tile_begin) * TILE + 4 * (int)threadIdx.x;
#pragma unroll
            for (int m = 0; m < Q; ++m) {
                const int c = c0 + 4 * m * BLOCK;
                ok[m] = c < T.n;
                const int cc = ok[m] ? c : T.n - 4;
                const CellIdx id = decompose(cc, T);
                const long o = T.off + cc;
                const Nb4 nb = neighbours4(id, T, o);
                oc[m] = o;
                mW[m] = nb.hW;
                mB[m] = nb.hB;
                const long idx[5] = {o, nb.W, nb.E, nb.S, nb.N};
#pragma unroll
                for (int v = 0; v < 5; ++v) {
                    zq[m][v] = ld4(z + idx[v]);
                    pq[m][v] = ld4(pold + idx[v]);
                }
                zb[m][0] = z[nb.B];
                pb[m][0] = pold[nb.B];
                zb[m][1] = z[nb.T];
                pb[m][1] = pold[nb.T];
                wq[m][0] = W::load4(a.w[0], a.scale, o);      // +r faces
                wq[m][1] = W::load4(a.w[0], a.scale, nb.W);   // -r faces (the W quad's +r)
                wq[m][2] = W::load4(a.w[1], a.scale, o);      // +theta faces
                wq[m][3] = W::load4(a.w[1], a.scale, nb.S);   // -theta faces
                wq[m][4] = W::load4(a.w[2], a.scale, o);      // +z faces
                wB[m] = W::load(a.w[2], a.scale, nb.B);       // -z face of the first cell
                ev[m] = W::load4(a.e, a.scale, o);
            }
        };
        MxDyn P;
        if (a.state_first) {   // a batch: converged systems load nothing
            P = a.dyn[1][s];
            if (!(P.done || P.broken)) load(t);
        } else {
            load(t);
            P = a.dyn[1][s];
        }
        MxDyn N = P;
        bool skip = P.done || P.broken;
        if (!skip) {
            double tot[2];
            reduce_tiles<2>(a.partB, a.line ? T.lt_begin : T.tile_begin, a.line ? T.lt_end : T.tile_end, tot, sh);
            if (!(tot[1] > 0.0)) {
                skip = true;
                N.broken = 1;
            } else {
                N.beta = P.fresh ? 0.0 : tot[1] / P.rz;
                N.rz = tot[1];
                N.fresh = 0;
            }
        }
        if (t == T.tile_begin && threadIdx.x == 0) store_state(&a.dyn[0][s], N);
        if (!skip) {
            const float fb = (float)N.beta;
            double acc[1] = {0.0};
            for (int u = 0; u < nrun; ++u) {
                if (u) load(t + u * w.stride);
#pragma unroll
                for (int m = 0; m < Q; ++m) {
                    float p[5][4];
#pragma unroll
                    for (int v = 0; v < 5; ++v)
#pragma unroll
                        for (int q = 0; q < 4; ++q) p[v][q] = zq[m][v].v[q] + fb * pq[m][v].v[q];
                    const float pB0 = zb[m][0] + fb * pb[m][0];
                    const float pT3 = zb[m][1] + fb * pb[m][1];
                    F4 y;
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const float pc = p[0][q];
                        const float pT = q == 3 ? pT3 : p[0][q + 1];
                        const float pB = q == 0 ? pB0 : p[0][q - 1];
                        const float wb = q == 0 ? (mB[m] ? wB[m] : 0.f) : wq[m][4].v[q - 1];
                        float yy = ev[m].v[q] * pc;
                        yy += wq[m][0].v[q] * (pc - p[2][q]);
                        yy += (mW[m] ? wq[m][1].v[q] : 0.f) * (pc - p[1][q]);
                        yy += wq[m][2].v[q] * (pc - p[4][q]);
                        yy += wq[m][3].v[q] * (pc - p[3][q]);
                        yy += wq[m][4].v[q] * (pc - pT);
                        yy += wb * (pc - pB);
                        y.v[q] = yy;
                    }
                    if (ok[m]) {
                        F4 pn;
#pragma unroll
                        for (int q = 0; q < 4; ++q) pn.v[q] = p[0][q];
                        st4(pnew + oc[m], pn);
                        st4(qo + oc[m], y);
#pragma unroll
                        for (int q = 0; q < 4; ++q) acc[0] += (double)p[0][q] * (double)y.v[q];
                    }
                }
            }
            block_sum<1>(acc, sh);
            store_run<1>(a.partA, t, nrun, w.stride, acc);
        }
        t += nrun * w.stride;
    }
}

template <int Q>
__global__ __launch_bounds__(BLOCK) void k_mupdate4(MxArgs a, int par) {
    constexpr int TILE = BLOCK * 4 * Q;
    __shared__ double sh[2 * (BLOCK / 64)];
    const float* __restrict__ pv = a.p[par ^ 1];
    float* __restrict__ d = a.d;
    float* __restrict__ r = a.r;
    float* __restrict__ zo = a.z;
    const float* __restrict__ q = a.q;
    const float* __restrict__ dinv = a.dinv;
    const TileWalk w = my_tiles(a.ntiles);
    for (int t = w.first; t < w.end;) {
        const TileDesc T = a.tiles[t];
        const int s = T.sys;
        const int nrun = w.run(t, T.tile_end);
        F4 dv[Q], pvv[Q], rv[Q], qv[Q], iv[Q];
        long oc[Q];
        bool ok[Q];
        auto load = [&](int tile) {
            const int c0 = (tile - T.tile_begin) * TILE + 4 * (int)threadIdx.x;
#pragma unroll
            for (int m = 0; m < Q; ++m) {
                const int c = c0 + 4 * m * BLOCK;
                ok[m] = c < T.n;
                const long o = T.off + (ok[m] ? c : T.n - 4);
                oc[m] = o;
                dv[m] = ld4nt(d + o);
                pvv[m] = ld4nt(pv + o);
                rv[m] = ld4nt(r + o);
                qv[m] = ld4nt(q + o);
                iv[m] = ld4nt(dinv + o);
            }
        };
        MxDyn P;
        if (a.state_first) {   // a batch: converged systems load nothing
            P = a.dyn[0][s];
            if (!(P.done || P.broken)) load(t);
        } else {
            load(t);
            P = a.dyn[0][s];
        }
        MxDyn N = P;
        bool skip = P.done || P.broken;
        if (!skip) {
            double tot[1];
            reduce_tiles<1>(a.partA, T.tile_begin, T.tile_begin + T.pa_n, tot, sh);
            if (!(fabs(tot[0]) > SMALL)) {
                skip = true;
                N.broken = 1;
            } else {
                N.alpha = P.rz / tot[0];
                N.total_inner = P.total_inner + 1;
            }
        }
        if (t == T.tile_begin && threadIdx.x == 0) store_state(&a.dyn[1][s], N);
        if (!skip) {
            const float fa = (float)N.alpha;
            double acc[2] = {0.0, 0.0};
            for (int u = 0; u < nrun; ++u) {
                if (u) load(t + u * w.stride);
#pragma unroll
                for (int m = 0; m < Q; ++m) {
                    F4 dn, rn, zn;
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        const float rr = rv[m].v[k] - fa * qv[m].v[k];
                        rn.v[k] = rr;
                        zn.v[k] = iv[m].v[k] * rr;
                        dn.v[k] = dv[m].v[k] + fa * pvv[m].v[k];
                    }
                    if (ok[m]) {
                        st4nt(d + oc[m], dn);
                        st4nt(r + oc[m], rn);
                        st4(zo + oc[m], zn);
#pragma unroll
                        for (int k = 0; k < 4; ++k) {
                            acc[0] += (double)rn.v[k] * (double)rn.v[k];
                            acc[1] += (double)rn.v[k] * (double)zn.v[k];
                        }
                    }
                }
            }
            block_sum<2>(acc, sh);
            store_run<2>(a.partB, t, nrun, w.stride, acc);
        }
        t += nrun * w.stride;
    }
}

// ---------------------------------------------------------------------------
// Weights of one quad as loaded, decoded only where they are used: a weight
// converted right after its load makes the wave wait for that load there.
// ---------------------------------------------------------------------------
template <int F>
struct WRaw;
template <>
struct WRaw<EAFP_COEF_FP32> {
    F4 v;
};
template <>
struct WRaw<EAFP_COEF_BF16> {
    unsigned x, y;   // two BF16 each, low half first
};
template <>
struct WRaw<EAFP_COEF_FP16> {
    unsigned x, y;   // two FP16 each
    F4 s;            // the row scales
};

template <int F>
__device__ __forceinline__ WRaw<F> wraw(const void* w, const float* sc, long o);
template <>
__device__ __forceinline__ WRaw<EAFP_COEF_FP32> wraw<EAFP_COEF_FP32>(const void* w, const float*, long o) {
    return {ld4((const float*)w + o)};
}
template <>
__device__ __forceinline__ WRaw<EAFP_COEF_BF16> wraw<EAFP_COEF_BF16>(const void* w, const float*, long o) {
    WRaw<EAFP_COEF_BF16> r;
    __builtin_memcpy(&r, __builtin_assume_aligned((const unsigned short*)w + o, 8), 8);
    return r;
}
template <>
__device__ __forceinline__ WRaw<EAFP_COEF_FP16> wraw<EAFP_COEF_FP16>(const void* w, const float* sc, long o) {
    WRaw<EAFP_COEF_FP16> r;
    __builtin_memcpy(&r, __builtin_assume_aligned((const _Float16*)w + o, 8), 8);
    r.s = ld4(sc + o);
    return r;
}

// The same values as WFmt<F>::load4 at the same cells.
__device__ __forceinline__ F4 wdec(const WRaw<EAFP_COEF_FP32>& r) { return r.v; }
__device__ __forceinline__ F4 wdec(const WRaw<EAFP_COEF_BF16>& r) {
    F4 f;
    f.v[0] = __uint_as_float(r.x << 16);
    f.v[1] = __uint_as_float(r.x & 0xffff0000u);
    f.v[2] = __uint_as_float(r.y << 16);
    f.v[3] = __uint_as_float(r.y & 0xffff0000u);
    return f;
}
__device__ __forceinline__ F4 wdec(const WRaw<EAFP_COEF_FP16>& r) {
    _Float16 h[4];
    __builtin_memcpy(h, &r, 8);
    F4 f = r.s;
#pragma unroll
    for (int q = 0; q < 4; ++q) f.v[q] *= (float)h[q];
    return f;
}

// The loads of one step of the column march (centre plane i): z and p_old
// of the plane beyond it (i+dir), its weights (w0 of its outer face, w1, w2,
// e) and its theta halo rows (z, p_old, w1 of row -1).
template <int F>
struct ColSet {
    F4 zE, oE, zH, oH;
    WRaw<F> r0, r1, r2, re, r1H;
};

// ---------------------------------------------------------------------------
// Column-marching form of k_mspmv (same state logic, same arithmetic per
// cell).  A block takes one column segment (ColDesc, which carries the
// system's geometry): nj theta-rows of one system over the r-planes
// i0 .. i1-1; each thread holds one quad (4 z-cells of a row).  p = z +
// beta*p_old of the planes i-1, i and i+1 stays in registers, so the
// r-neighbours are loaded once, as the plane itself; p and the +theta weight
// of plane i go to a double-buffered LDS slab (with the two theta halo rows,
// loaded by the first 2*nz/4 threads), from which the theta and z
// neighbours are read: one barrier per plane.  Per cell it reads z, p_old,
// the 3 weights and e once, plus the halo rows (z and p_old of 2 rows, w1 of
// one, per nj rows), and writes p and q.  A segment marches up in r
// (C.dir = +1) or down (-1): the segments of a column alternate, so the two
// segments on either side of a boundary read the planes there (each one's
// extra plane beyond its range) at about the same time, and the second read
// finds them in L2.
// Pipelining: every load is issued one whole step before its use and no
// loaded value is carried across the loop's back-edge.  The march is
// unrolled by two over two ColSets: step t first issues the loads of step
// t+1 into one set, then works from the other, loaded during step t-1.  The
// weights stay raw until after the plane's barrier, and every lane issues
// the same memory instructions (lanes past the slab store to the sink past
// the arenas), so the waits of a step never cover the loads it just issued.
// ---------------------------------------------------------------------------
template <int F>
__global__ __launch_bounds__(BLOCK) void k_mspmv_col(MxArgs a, int par) {
    constexpr int SP = COL_CELLS + 2 * COL_MAX_NZ;   // p slab: rows -1 .. nj
    constexpr int SW = COL_CELLS + COL_MAX_NZ;       // w1 slab: rows -1 .. nj-1
    __shared__ double sh[2 * (BLOCK / 64)];
    __shared__ float Sp[2][SP];
    __shared__ float Sw1[2][SW];
    __shared__ float Sw2[2][BLOCK];
    const int G = gridDim.x;
    const ColDesc C = a.cols[(blockIdx.x % NXCD) * (G / NXCD) + blockIdx.x / NXCD];
    if (C.tile < 0) return;
    const int s = C.sys;
    const float* __restrict__ pold = a.p[par];
    float* __restrict__ pnew = a.p[par ^ 1];
    const float* __restrict__ z = a.z;
    float* __restrict__ qo = a.q;
    const int nz = C.nz, nth = C.nth, nr = C.nr;
    const int Qr = nz >> 2;                       // quads per row
    const int tid = threadIdx.x;
    const int row = tid / Qr, kq = tid - row * Qr;
    const bool act = row < C.nj;
    // a lane past the slab shadows row 0: valid addresses, results unused
    const long qoff = (long)(C.j0 + (act ? row : 0)) * nz + 4 * (act ? kq : 0);
    const bool hal = tid < 2 * Qr;                // halo lanes: row -1, then row nj
    const int hrow = tid < Qr ? -1 : C.nj;
    const int hk = hal ? 4 * (tid < Qr ? tid : tid - Qr) : 0;
    const long hoff = (long)(((C.j0 + hrow) % nth + nth) % nth) * nz + hk;
    const long sr = (long)nth * nz;
    // a plane outside 0 .. nr-1 is clamped to the nearest one: its loads stay
    // in the system, and as a neighbour it equals the cell's own p (face term 0)
    auto base = [&](int i) { return C.off + (long)(i < 0 ? 0 : i < nr ? i : nr - 1) * sr; };
    const int dir = C.dir;
    const int i0 = dir > 0 ? C.i0 : C.i1 - 1;   // first plane of the march
    const int np = C.i1 - C.i0;

    auto issue = [&](int i, ColSet<F>& S) {
        S.zE = ld4(z + base(i + dir) + qoff);
        S.oE = ld4(pold + base(i + dir) + qoff);
        S.r0 = wraw<F>(a.w[0], a.scale, base(dir > 0 ? i : i - 1) + qoff);   // the face towards i+dir
        S.r1 = wraw<F>(a.w[1], a.scale, base(i) + qoff);
        S.r2 = wraw<F>(a.w[2], a.scale, base(i) + qoff);
        S.re = wraw<F>(a.e, a.scale, base(i) + qoff);
        S.zH = ld4(z + base(i) + hoff);
        S.oH = ld4(pold + base(i) + hoff);
        S.r1H = wraw<F>(a.w[1], a.scale, base(i) + hoff);
    };
    // planes i0-dir and i0 of z and p_old, the face between them, and the
    // first step's set: in flight while the partials are reduced (in a
    // batch, only once the system's state shows it still running)
    F4 zP, oP, zC, oC;
    WRaw<F> rIn;   // the face between the centre plane and the one before it in the march
    ColSet<F> A, B;
    auto prologue = [&]() {
        zP = ld4(z + base(i0 - dir) + qoff);
        oP = ld4(pold + base(i0 - dir) + qoff);
        zC = ld4(z + base(i0) + qoff);
        oC = ld4(pold + base(i0) + qoff);
        rIn = wraw<F>(a.w[0], a.scale, base(dir > 0 ? i0 - 1 : i0) + qoff);
        issue(i0, A);
    };
    // the partials' loads go out ahead of the cell loads, so they return first
    const int pb0 = a.line ? C.lt_begin : C.tile, pb1 = a.line ? C.lt_end : C.tile_end;
    TilePre<2> pre;
    MxDyn P;
    if (a.state_first) {
        P = a.dyn[1][s];
        if (!(P.done || P.broken)) {
            reduce_tiles_load<2>(a.partB, pb0, pb1, pre);
            prologue();
        }
    } else {
        reduce_tiles_load<2>(a.partB, pb0, pb1, pre);
        prologue();
        P = a.dyn[1][s];
    }
    MxDyn N = P;
    bool skip = P.done || P.broken;
    if (!skip) {
        double tot[2];
        reduce_tiles_sum<2>(a.partB, pb0, pb1, pre, tot, sh);
        if (!(tot[1] > 0.0)) {
            skip = true;
            N.broken = 1;
        } else {
            N.beta = P.fresh ? 0.0 : tot[1] / P.rz;
            N.rz = tot[1];
            N.fresh = 0;
        }
    }
    if (C.seg == 0 && tid == 0) store_state(&a.dyn[0][s], N);
    if (skip) return;
    const float fb = (float)N.beta;
    auto pv = [&](const F4& zz, const F4& oo) {
        F4 r;
#pragma unroll
        for (int q = 0; q < 4; ++q) r.v[q] = zz.v[q] + fb * oo.v[q];
        return r;
    };
    F4 pW = pv(zP, oP), pC = pv(zC, oC);   // W: the plane before the centre in the march
    double acc[1] = {0.0};
    auto step = [&](int t, ColSet<F>& cur, ColSet<F>& nxt) {
        const int i = i0 + dir * t;
        const int b = t & 1;
        issue(t + 1 < np ? i + dir : i, nxt);   // past the last plane: a reload, unused
        const F4 pE = pv(cur.zE, cur.oE);
        const F4 w1 = wdec(cur.r1), w2 = wdec(cur.r2);
        // plane i into the slab: p (rows -1 .. nj), w1 (rows -1 .. nj-1), w2 of each quad's last cell
        if (act) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                Sp[b][(row + 1) * nz + 4 * kq + q] = pC.v[q];
                Sw1[b][(row + 1) * nz + 4 * kq + q] = w1.v[q];
            }
        }
        if (hal) {
            const F4 pH = pv(cur.zH, cur.oH);
#pragma unroll
            for (int q = 0; q < 4; ++q) Sp[b][(hrow + 1) * nz + hk + q] = pH.v[q];
            if (hrow < 0) {
                const F4 w1H = wdec(cur.r1H);
#pragma unroll
                for (int q = 0; q < 4; ++q) Sw1[b][hk + q] = w1H.v[q];
            }
        }
        Sw2[b][tid] = w2.v[3];
        __syncthreads();
        // +r face of the cell: w0 of plane i; -r face: w0 of plane i-1
        const F4 w0 = wdec(dir > 0 ? cur.r0 : rIn), w0W = wdec(dir > 0 ? rIn : cur.r0), ev = wdec(cur.re);
        const F4& pRE = dir > 0 ? pE : pW;   // p of plane i+1
        const F4& pRW = dir > 0 ? pW : pE;   // p of plane i-1
        const float* sp = &Sp[b][(row + 1) * nz + 4 * kq];
        const float* sw = &Sw1[b][row * nz + 4 * kq];
        const bool hW = i > 0;
        const float pB0 = kq > 0 ? sp[-1] : pC.v[0];
        const float pT3 = kq < Qr - 1 ? sp[4] : pC.v[3];
        const float wB0 = kq > 0 ? Sw2[b][tid - 1] : 0.f;
        F4 y;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float pc = pC.v[q];
            const float pS = sp[q - nz], pN = sp[q + nz];
            const float pT = q == 3 ? pT3 : pC.v[q + 1];
            const float pB = q == 0 ? pB0 : pC.v[q - 1];
            const float wb = q == 0 ? wB0 : w2.v[q - 1];
            float yy = ev.v[q] * pc;
            yy += w0.v[q] * (pc - pRE.v[q]);
            yy += (hW ? w0W.v[q] : 0.f) * (pc - pRW.v[q]);
            yy += w1.v[q] * (pc - pN);
            yy += sw[q] * (pc - pS);
            yy += w2.v[q] * (pc - pT);
            yy += wb * (pc - pB);
            y.v[q] = yy;
        }
        const long o = act ? base(i) + qoff : a.sink;
        st4(pnew + o, pC);
        st4(qo + o, y);
        if (act) {
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[0] += (double)pC.v[q] * (double)y.v[q];
        }
        rIn = cur.r0;   // the face between i and i+dir: the next centre's inner face
        pW = pC;
        pC = pE;
    };
    int t = 0;
    for (; t + 1 < np; t += 2) {
        step(t, A, B);
        step(t + 1, B, A);
    }
    if (t < np) step(t, A, B);
    block_sum<1>(acc, sh);
    if (tid == 0) a.partA[(long)C.tile + C.seg] = acc[0];   // the consumers read nseg slots
}

// ---------------------------------------------------------------------------
// Column-marching stencil for any nz (at most COL_MAX_NZ): as k_mspmv_col,
// alternating march directions included, but a thread takes the slab cells
// tid + m*BLOCK (m < 4) of each plane with
// 4-byte accesses, and the z neighbours of every cell come from the LDS
// slab too (w+z of the whole slab is staged for the -z faces).
// ---------------------------------------------------------------------------
template <int F>
__global__ __launch_bounds__(BLOCK) void k_mspmv_colg(MxArgs a, int par) {
    using W = WFmt<F>;
    constexpr int M = COL_CELLS / BLOCK;             // slab cells per thread
    constexpr int SP = COL_CELLS + 2 * COL_MAX_NZ;   // p slab: rows -1 .. nj
    constexpr int SW = COL_CELLS + COL_MAX_NZ;       // w+theta slab: rows -1 .. nj-1
    __shared__ double sh[2 * (BLOCK / 64)];
    __shared__ float Sp[2][SP];
    __shared__ float Sw1[2][SW];
    __shared__ float Sw2[2][COL_CELLS];
    const int G = gridDim.x;
    const ColDesc C = a.cols[(blockIdx.x % NXCD) * (G / NXCD) + blockIdx.x / NXCD];
    if (C.tile < 0) return;
    const TileDesc T = a.tiles[C.tile];
    const int s = T.sys;
    const float* __restrict__ pold = a.p[par];
    float* __restrict__ pnew = a.p[par ^ 1];
    const float* __restrict__ z = a.z;
    float* __restrict__ qo = a.q;
    const int nz = T.nz, nth = T.nth, nr = T.nr;
    const int tid = threadIdx.x;
    const int ncell = C.nj * nz;
    int row[M], kk[M];
    long coff[M];
    bool act[M];
#pragma unroll
    for (int m = 0; m < M; ++m) {
        const int c = tid + m * BLOCK;
        act[m] = c < ncell;
        const int cc = act[m] ? c : 0;               // lanes past the slab shadow cell 0
        row[m] = fdiv(cc, nz, T.inv_nz);
        kk[m] = cc - row[m] * nz;
        coff[m] = (long)(C.j0 + row[m]) * nz + kk[m];
    }
    // halo cells h = tid + m*BLOCK < 2*nz: row -1 (h < nz), then row nj
    constexpr int MH = 2 * COL_MAX_NZ / BLOCK;
    bool hal[MH];
    int hslot[MH];
    long hoff[MH];
#pragma unroll
    for (int m = 0; m < MH; ++m) {
        const int h = tid + m * BLOCK;
        hal[m] = h < 2 * nz;
        const int hc = hal[m] ? h : 0;
        const int hrow = hc < nz ? -1 : C.nj;
        const int hk = hc < nz ? hc : hc - nz;
        hslot[m] = (hrow + 1) * nz + hk;
        hoff[m] = (long)(((C.j0 + hrow) % nth + nth) % nth) * nz + hk;
    }
    const long sr = (long)nth * nz;
    auto base = [&](int i) { return T.off + (long)(i < 0 ? 0 : i < nr ? i : nr - 1) * sr; };
    const int dir = C.dir;                      // as in k_mspmv_col
    const int i0 = dir > 0 ? C.i0 : C.i1 - 1;   // first plane of the march
    const int np = C.i1 - C.i0;
    const long bW = base(i0 - 1), b0 = base(i0), b1 = base(i0 + 1);
    float zW[M], oW[M], w0W[M], zC[M], oC[M], zE[M], oE[M], w0[M], w1[M], w2[M], ev[M];
    float zH[MH], oH[MH], w1H[MH];
    auto prologue = [&]() {
#pragma unroll
        for (int m = 0; m < M; ++m) {
            zW[m] = z[bW + coff[m]];
            oW[m] = pold[bW + coff[m]];
            w0W[m] = W::load(a.w[0], a.scale, bW + coff[m]);
            zC[m] = z[b0 + coff[m]];
            oC[m] = pold[b0 + coff[m]];
            zE[m] = z[b1 + coff[m]];
            oE[m] = pold[b1 + coff[m]];
            w0[m] = W::load(a.w[0], a.scale, b0 + coff[m]);
            w1[m] = W::load(a.w[1], a.scale, b0 + coff[m]);
            w2[m] = W::load(a.w[2], a.scale, b0 + coff[m]);
            ev[m] = W::load(a.e, a.scale, b0 + coff[m]);
        }
#pragma unroll
        for (int m = 0; m < MH; ++m) {
            zH[m] = z[b0 + hoff[m]];
            oH[m] = pold[b0 + hoff[m]];
            w1H[m] = W::load(a.w[1], a.scale, b0 + hoff[m]);
        }
    };
    MxDyn P;
    if (a.state_first) {   // a batch: converged systems load nothing
        P = a.dyn[1][s];
        if (!(P.done || P.broken)) prologue();
    } else {
        prologue();
        P = a.dyn[1][s];
    }
    MxDyn N = P;
    bool skip = P.done || P.broken;
    if (!skip) {
        double tot[2];
        reduce_tiles<2>(a.partB, a.line ? T.lt_begin : T.tile_begin, a.line ? T.lt_end : T.tile_end, tot, sh);
        if (!(tot[1] > 0.0)) {
            skip = true;
            N.broken = 1;
        } else {
            N.beta = P.fresh ? 0.0 : tot[1] / P.rz;
            N.rz = tot[1];
            N.fresh = 0;
        }
    }
    if (C.seg == 0 && tid == 0) store_state(&a.dyn[0][s], N);
    if (skip) return;
    const float fb = (float)N.beta;
    float pW[M], pC[M], pE[M], pH[MH];
#pragma unroll
    for (int m = 0; m < M; ++m) {
        pW[m] = zW[m] + fb * oW[m];
        pC[m] = zC[m] + fb * oC[m];
        pE[m] = zE[m] + fb * oE[m];
    }
#pragma unroll
    for (int m = 0; m < MH; ++m) pH[m] = zH[m] + fb * oH[m];
    double acc[1] = {0.0};
    for (int t = 0; t < np; ++t) {
        const int i = i0 + dir * t;
        const int b = i & 1;
        const bool more = t + 1 < np;
        float zN2[M], oN2[M], w0n[M], w1n[M], w2n[M], evn[M], zHn[MH], oHn[MH], w1Hn[MH];
        if (more) {
            const int in = i + dir;
            const long b2 = base(i + 2 * dir), bn = base(in), bf = base(dir > 0 ? in : in - 1);
#pragma unroll
            for (int m = 0; m < M; ++m) {
                zN2[m] = z[b2 + coff[m]];
                oN2[m] = pold[b2 + coff[m]];
                w0n[m] = W::load(a.w[0], a.scale, bf + coff[m]);
                w1n[m] = W::load(a.w[1], a.scale, bn + coff[m]);
                w2n[m] = W::load(a.w[2], a.scale, bn + coff[m]);
                evn[m] = W::load(a.e, a.scale, bn + coff[m]);
            }
#pragma unroll
            for (int m = 0; m < MH; ++m) {
                zHn[m] = z[bn + hoff[m]];
                oHn[m] = pold[bn + hoff[m]];
                w1Hn[m] = W::load(a.w[1], a.scale, bn + hoff[m]);
            }
        }
#pragma unroll
        for (int m = 0; m < M; ++m) {
            if (act[m]) {
                Sp[b][(row[m] + 1) * nz + kk[m]] = pC[m];
                Sw1[b][(row[m] + 1) * nz + kk[m]] = w1[m];
                Sw2[b][row[m] * nz + kk[m]] = w2[m];
            }
        }
#pragma unroll
        for (int m = 0; m < MH; ++m) {
            if (hal[m]) {
                Sp[b][hslot[m]] = pH[m];
                if (hslot[m] < nz) Sw1[b][hslot[m]] = w1H[m];
            }
        }
        __syncthreads();
        const bool hW = i > 0;
        const long bi = base(i);
#pragma unroll
        for (int m = 0; m < M; ++m) {
            if (act[m]) {
                const int sl = (row[m] + 1) * nz + kk[m];
                const float pc = pC[m];
                const bool hB = kk[m] > 0, hT = kk[m] < nz - 1;
                const float pS = Sp[b][sl - nz], pN = Sp[b][sl + nz];
                const float pB = hB ? Sp[b][sl - 1] : pc;
                const float pT = hT ? Sp[b][sl + 1] : pc;
                const float wS = Sw1[b][sl - nz];
                const float wB = hB ? Sw2[b][row[m] * nz + kk[m] - 1] : 0.f;
                float y = ev[m] * pc;
                y += w0[m] * (pc - pE[m]);
                y += (hW ? w0W[m] : 0.f) * (pc - pW[m]);
                y += w1[m] * (pc - pN);
                y += wS * (pc - pS);
                y += w2[m] * (pc - pT);
                y += wB * (pc - pB);
                pnew[bi + coff[m]] = pc;
                qo[bi + coff[m]] = y;
                acc[0] += (double)pc * (double)y;
            }
        }
        if (more) {
#pragma unroll
            for (int m = 0; m < M; ++m) {
                const float pn = zN2[m] + fb * oN2[m];
                if (dir > 0) {
                    pW[m] = pC[m];
                    pC[m] = pE[m];
                    pE[m] = pn;
                    w0W[m] = w0[m];
                    w0[m] = w0n[m];
                } else {
                    pE[m] = pC[m];
                    pC[m] = pW[m];
                    pW[m] = pn;
                    w0[m] = w0W[m];
                    w0W[m] = w0n[m];
                }
                w1[m] = w1n[m];
                w2[m] = w2n[m];
                ev[m] = evn[m];
            }
#pragma unroll
            for (int m = 0; m < MH; ++m) {
                pH[m] = zHn[m] + fb * oHn[m];
                w1H[m] = w1Hn[m];
            }
        }
    }
    block_sum<1>(acc, sh);
    if (tid == 0) a.partA[(long)T.tile_begin + C.seg] = acc[0];   // the consumers read nseg slots
}

// Finished systems in slot 1 (one block).
// ---------------------------------------------------------------------------
// z-line preconditioner.  M is the tridiagonal part along z of the rounded
// operator (diagonal D as the Jacobi inverse sees it, off-diagonals minus
// the +z face weights), factored M = L diag(d) L^T with L unit lower
// bidiagonal: l_k = -w_{k-1}/d_{k-1}, d_k = D_k - w_{k-1}^2/d_{k-1}.  Any
// positive d gives a symmetric positive definite M, so d_k is kept at or
// above D_k/16 on nearly singular lines (an isolated Neumann column), which
// bounds the amplification of z = M^-1 r.  Rows with D = 0 get l = 0,
// 1/d = 0 (z = 0, as Jacobi's 1/D = 0).
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned pack_fac(double l, double inv) {
    const _Float16 lh = (_Float16)(float)l;
    unsigned short lb;
    __builtin_memcpy(&lb, &lh, 2);
    return (unsigned)lb | ((unsigned)f32_to_bf16_rne((float)inv) << 16);
}

__device__ __forceinline__ void unpack_fac(float bits, float& l, float& inv) {
    const unsigned u = __float_as_uint(bits);
    const unsigned short lb = (unsigned short)(u & 0xFFFFu);
    _Float16 lh;
    __builtin_memcpy(&lh, &lb, 2);
    l = (float)lh;
    inv = __uint_as_float(u & 0xFFFF0000u);
}

// One thread per line, k = 0 .. nz-1 in order (once per solve, after
// k_compress has left D in q).
template <int F>
__global__ __launch_bounds__(BLOCK) void k_linefac(MxArgs a) {
    using W = WFmt<F>;
    const LineTile L = a.lt[blockIdx.x];
    const int li = threadIdx.x;
    if (li >= L.nlines) return;
    const long o0 = L.off + (long)(L.line0 + li) * L.nz;
    double dprev = 0.0, wprev = 0.0;
    for (int k = 0; k < L.nz; ++k) {
        const long o = o0 + k;
        const double Dk = a.q[o];
        double l = 0.0, dk = Dk;
        if (k > 0 && dprev > 0.0) {
            l = -wprev / dprev;
            dk = Dk - wprev * wprev / dprev;
        }
        dk = fmax(dk, 0.0625 * Dk);
        wprev = W::load(a.w[2], a.scale, o);
        const bool live = Dk > 0.0;
        a.fac[o] = pack_fac(live ? l : 0.0, live ? 1.0 / dk : 0.0);
        dprev = live ? dk : 0.0;
    }
}

// z = M^-1 r on the block's lines: thread t holds 4 consecutive cells of
// line li = t/nq (lane j = t%nq).  Forward L y = r and backward
// L^T z = y/d are first-order recurrences; every lane composes its 4
// cells into one affine map, the maps go through LDS, and each lane folds
// the maps of the lanes before (forward) or after (backward) it in its line.
// Called by the whole block.
struct LineLds {
    float2 fw[BLOCK];
    float2 bw[BLOCK];
    float lf[BLOCK];
};

__device__ __forceinline__ void line_solve(LineLds& S, bool inl, int li, int j, int nq, const float (&r)[4],
                                           const float (&l)[4], const float (&iv)[4], float (&z)[4]) {
    const int tid = threadIdx.x;
    float A = 1.f, Bv = 0.f;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        A = -l[k] * A;
        Bv = r[k] - l[k] * Bv;
    }
    S.fw[tid] = make_float2(A, Bv);
    S.lf[tid] = l[0];
    __syncthreads();
    // the folds read 8 maps ahead of the multiply-add chain (same order of
    // operations as one map at a time)
    float y = 0.f;
    if (inl) {
        const float2* f = &S.fw[li * nq];
        int m = 0;
        for (; m + 8 <= j; m += 8) {
            float2 t[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) t[u] = f[m + u];
#pragma unroll
            for (int u = 0; u < 8; ++u) y = t[u].x * y + t[u].y;
        }
        for (; m < j; ++m) {
            const float2 t = f[m];
            y = t.x * y + t.y;
        }
    }
    float u[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        y = r[k] - l[k] * y;
        u[k] = y * iv[k];
    }
    const float ln = (inl && j + 1 < nq) ? S.lf[tid + 1] : 0.f;
    float Ab = -ln, Bb = u[3];
#pragma unroll
    for (int k = 2; k >= 0; --k) {
        Ab = -l[k + 1] * Ab;
        Bb = u[k] - l[k + 1] * Bb;
    }
    S.bw[tid] = make_float2(Ab, Bb);
    __syncthreads();
    float zin = 0.f;
    if (inl) {
        const float2* f = &S.bw[li * nq];
        int m = nq - 1;
        for (; m - 8 >= j; m -= 8) {
            float2 t[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) t[u] = f[m - u];
#pragma unroll
            for (int u = 0; u < 8; ++u) zin = t[u].x * zin + t[u].y;
        }
        for (; m > j; --m) {
            const float2 t = f[m];
            zin = t.x * zin + t.y;
        }
    }
    z[3] = u[3] - ln * zin;
#pragma unroll
    for (int k = 2; k >= 0; --k) z[k] = u[k] - l[k + 1] * z[k + 1];
}

// The cells of a line-tile thread.  Quad form (every nz of the batch a
// multiple of 4): cells o .. o+3, one 16-byte access each.  Generic form:
// the block's lines are one run of n = nlines*nz cells starting at `run`;
// the block moves every array through LDS in 16-byte chunks aligned to the
// arena (stage/unstage; the chunks cover [run & ~3, run + n) rounded up to
// 4, inside the system's 32-cell-aligned, padded slot of the arena), and
// the thread takes its cells li*nz + 4j .. +nv-1 of the run from there
// (nv < 4 only in a line's last lane; slots past nv repeat its first
// cell).  Unstaging writes only the run's cells.  Threads past the tile's
// lines take the run's first cell.
constexpr int LINE_RUN = 4 * BLOCK;   // cells of a line tile at most
constexpr int LINE_BUF = LINE_RUN + 4; // LDS floats per staged array (alignment slack)

template <bool Q>
struct LineCells {
    long o;      // quad form: first cell
    long run;    // generic form: first cell of the block's run
    int n;       // generic form: cells in the run
    int c0;      // generic form: the thread's first cell within the run
    int nv;
    bool inl;
    int li, j, nq;
    __device__ LineCells(const LineTile& L) {
        const int tid = threadIdx.x;
        nq = line_lanes(L.nz);
        li = tid / nq;
        j = tid - li * nq;
        inl = li < L.nlines;
        run = L.off + (long)L.line0 * L.nz;
        n = L.nlines * L.nz;
        if (Q) {
            o = run + 4 * (long)(inl ? tid : 0);
            nv = 4;
            c0 = 0;
        } else {
            o = run;
            c0 = inl ? li * L.nz + 4 * j : 0;
            nv = inl ? min(4, L.nz - 4 * j) : 1;
        }
    }
    // generic form: buf[m] holds cell (run & ~3) + m; the run's cell c is buf[sh + c]
    __device__ int shift() const { return (int)(run & 3); }
    __device__ int chunks() const { return (shift() + n + 3) >> 2; }
    // the run of p into buf (the caller syncs before reading)
    __device__ void stage(const float* p, float* buf) const {
        const float* g = p + (run - shift());
        for (int q = threadIdx.x; q < chunks(); q += BLOCK) {
            const F4 v = ld4nt(g + 4 * q);
            __builtin_memcpy(buf + 4 * q, &v, 16);
        }
    }
    __device__ F4 from(const float* buf) const {
        F4 r;
        const int b = shift() + c0;
#pragma unroll
        for (int k = 0; k < 4; ++k) r.v[k] = buf[b + (k < nv ? k : 0)];
        return r;
    }
    __device__ void to(float* buf, const F4& v) const {
        if (!inl) return;
        const int b = shift() + c0;
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (k < nv) buf[b + k] = v.v[k];
    }
    // buf back to the run's cells of p (the caller syncs before): whole
    // chunks with one 16-byte store, the run's partial first and last chunks
    // cell by cell
    __device__ void unstage(float* p, const float* buf, bool nt) const {
        const int sh = shift(), nc = chunks();
        float* g = p + (run - sh);
        for (int q = threadIdx.x; q < nc; q += BLOCK) {
            const int m0 = 4 * q;
            if (m0 >= sh && m0 + 4 <= sh + n) {
                F4 v;
                __builtin_memcpy(&v, buf + m0, 16);
                if (nt) st4nt(g + m0, v);
                else st4(g + m0, v);
            } else {
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int m = m0 + k;
                    if (m >= sh && m < sh + n) g[m] = buf[m];
                }
            }
        }
    }
};

// k_mupdate with the z-line preconditioner: one line tile per block.
template <bool Q>
__global__ __launch_bounds__(BLOCK) void k_mupdate_line(MxArgs a, int par) {
    __shared__ double sh[2 * (BLOCK / 64)];
    __shared__ LineLds S;
    __shared__ __align__(16) float G[Q ? 1 : 5][Q ? 4 : LINE_BUF];   // generic form: d, p, r, q, factors of the run
    const LineTile L = a.lt[blockIdx.x];
    const int s = L.sys;
    const LineCells<Q> C(L);
    F4 dv, pvv, rv, qv, fv;
    auto load_cells = [&]() {
        if (Q) {
            dv = ld4nt(a.d + C.o);
            pvv = ld4nt(a.p[par ^ 1] + C.o);
            rv = ld4nt(a.r + C.o);
            qv = ld4nt(a.q + C.o);
            fv = ld4nt((const float*)a.fac + C.o);
        } else {
            C.stage(a.d, G[0]);
            C.stage(a.p[par ^ 1], G[1]);
            C.stage(a.r, G[2]);
            C.stage(a.q, G[3]);
            C.stage((const float*)a.fac, G[4]);
        }
    };
    // one system: the partials' loads go out ahead of the cell loads (a
    // batch keeps the plain reduction after them: there the early loads
    // measured 1.7 us slower per launch)
    TilePre<1> pre;
    MxDyn P;
    if (a.state_first) {
        P = a.dyn[0][s];
        if (!(P.done || P.broken)) load_cells();
    } else {
        reduce_tiles_load<1>(a.partA, L.tile_begin, L.tile_begin + L.pa_n, pre);
        load_cells();
        P = a.dyn[0][s];
    }
    MxDyn N = P;
    bool skip = P.done || P.broken;
    if (!skip) {
        double tot[1];
        if (a.state_first) reduce_tiles<1>(a.partA, L.tile_begin, L.tile_begin + L.pa_n, tot, sh);   // syncs the block
        else reduce_tiles_sum<1>(a.partA, L.tile_begin, L.tile_begin + L.pa_n, pre, tot, sh);
        if (!(fabs(tot[0]) > SMALL)) {
            skip = true;
            N.broken = 1;
        } else {
            N.alpha = P.rz / tot[0];
            N.total_inner = P.total_inner + 1;
        }
    }
    if ((int)blockIdx.x == L.lt_begin && threadIdx.x == 0) store_state(&a.dyn[1][s], N);
    if (skip) return;   // uniform over the block
    if (!Q) {
        dv = C.from(G[0]);
        pvv = C.from(G[1]);
        rv = C.from(G[2]);
        qv = C.from(G[3]);
        fv = C.from(G[4]);
    }
    const float fa = (float)N.alpha;
    float rn[4], l[4], iv[4], zn[4];
    F4 dn, rr, zz;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const bool v = k < C.nv;
        rn[k] = v ? rv.v[k] - fa * qv.v[k] : 0.f;
        dn.v[k] = dv.v[k] + fa * pvv.v[k];
        unpack_fac(fv.v[k], l[k], iv[k]);
        l[k] = v ? l[k] : 0.f;
        iv[k] = v ? iv[k] : 0.f;
    }
    line_solve(S, C.inl, C.li, C.j, C.nq, rn, l, iv, zn);
    double acc[2] = {0.0, 0.0};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        rr.v[k] = rn[k];
        zz.v[k] = zn[k];
        if (C.inl) {
            acc[0] += (double)rn[k] * (double)rn[k];
            acc[1] += (double)rn[k] * (double)zn[k];
        }
    }
    if (Q) {
        if (C.inl) {
            st4nt(a.d + C.o, dn);
            st4nt(a.r + C.o, rr);
            st4(a.z + C.o, zz);
        }
    } else {
        // every thread has read its cells of G (line_solve synced since): reuse G[0..2]
        C.to(G[0], dn);
        C.to(G[1], rr);
        C.to(G[2], zz);
        __syncthreads();
        C.unstage(a.d, G[0], true);
        C.unstage(a.r, G[1], true);
        C.unstage(a.z, G[2], false);
    }
    block_sum<2>(acc, sh);
    store_run<2>(a.partB, (int)blockIdx.x, 1, 1, acc);
}

// k_mreplace with the z-line preconditioner: one line tile per block.
template <bool Q>
__global__ __launch_bounds__(BLOCK) void k_mreplace_line(MxArgs a) {
    __shared__ double sh[NRES * (BLOCK / 64)];
    __shared__ LineLds S;
    const LineTile L = a.lt[blockIdx.x];
    const int s = L.sys;
    const bool writer = (int)blockIdx.x == L.lt_begin && threadIdx.x == 0;
    const MxStat P = a.stat[0][s];
    bool done = P.done != 0;
    bool zero_p = false;
    double mean_r[MAX_COMP] = {};
    if (!done) {
        const MxConst& K = a.cst[s];
        double tot[NRES];
#pragma unroll
        for (int c = 0; c < NRES; ++c) tot[c] = a.sumR[(long)s * NRES + c];
        const int rnd = P.n_refine;
        MxStat N = P;
        N.bnorm = rnd == 0 ? sqrt(tot[2]) : P.bnorm;
        N.b1 = rnd == 0 ? tot[3] : P.b1;
        int status = EAFP_STOP_RUNNING;
        if (N.bnorm < SMALL) {
            N.res_l2 = 0.0;
            N.res_l1 = 0.0;
            status = EAFP_STOP_ZERO_RHS;
        } else {
            N.res_l2 = sqrt(tot[0]) / N.bnorm;
            N.res_l1 = N.b1 > 0.0 ? tot[1] / N.b1 : tot[1];
            if (N.res_l2 < K.tol) status = EAFP_STOP_CONVERGED;
            else if (rnd >= K.max_steps) status = EAFP_STOP_MAX_ITER;
        }
        done = status != EAFP_STOP_RUNNING;
        const MxDyn D0 = a.dyn[0][s];
        zero_p = (rnd == 0) || K.restart || D0.broken;
#pragma unroll
        for (int q = 0; q < MAX_COMP; ++q) {
            mean_r[q] = q < K.ncomp ? tot[4 + q] * K.inv_cnt[q] : 0.0;
            N.mean_x[q] = q < K.ncomp ? tot[4 + MAX_COMP + q] * K.inv_cnt[q] : 0.0;
        }
        N.xbuf = (rnd + 1) & 1;
        MxDyn D = D0;
        if (done) {
            N.done = 1;
            N.status = status;
            D.done = 1;
        } else {
            N.n_refine = rnd + 1;
            D.fresh = zero_p ? 1 : 0;
            D.broken = 0;
        }
        if (writer) {
            store_state(&a.stat[1][s], N);
            store_state(&a.dyn[1][s], D);
        }
    } else if (writer) {
        const MxDyn D0 = a.dyn[0][s];
        store_state(&a.stat[1][s], P);
        store_state(&a.dyn[1][s], D0);
    }
    if (done) return;   // uniform over the block
    // once per refinement step: the generic form reads and writes its cells directly
    const LineCells<Q> C(L);
    const long oc0 = Q ? C.o : C.run + C.c0;
    F4 fv;
    if (Q) {
        fv = ld4nt((const float*)a.fac + oc0);
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) fv.v[k] = ((const float*)a.fac)[oc0 + (k < C.nv ? k : 0)];
    }
    float rr[4], l[4], iv[4], zn[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const bool v = k < C.nv;
        const long oc = oc0 + (v ? k : 0);
        const int slot = a.comp[oc];
        const double rv = __builtin_nontemporal_load(a.r64 + oc);
        double mu = 0.0;
#pragma unroll
        for (int q = 0; q < MAX_COMP; ++q) mu = slot == q ? mean_r[q] : mu;
        rr[k] = v ? (float)(rv - mu) : 0.f;
        unpack_fac(fv.v[k], l[k], iv[k]);
        l[k] = v ? l[k] : 0.f;
        iv[k] = v ? iv[k] : 0.f;
    }
    line_solve(S, C.inl, C.li, C.j, C.nq, rr, l, iv, zn);
    double acc[2] = {0.0, 0.0};
    if (C.inl) {
        F4 rv4, zv4, zero;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            rv4.v[k] = rr[k];
            zv4.v[k] = zn[k];
            zero.v[k] = 0.0f;
            acc[0] += (double)rr[k] * (double)rr[k];
            acc[1] += (double)rr[k] * (double)zn[k];
        }
        if (Q) {
            st4nt(a.r + oc0, rv4);
            st4(a.z + oc0, zv4);
            st4nt(a.d + oc0, zero);
            if (zero_p) st4(a.p[0] + oc0, zero);
        } else {
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (k < C.nv) {
                    __builtin_nontemporal_store(rr[k], a.r + oc0 + k);
                    a.z[oc0 + k] = zn[k];
                    __builtin_nontemporal_store(0.0f, a.d + oc0 + k);
                    if (zero_p) a.p[0][oc0 + k] = 0.0f;
                }
        }
    }
    block_sum<2>(acc, sh);
    store_run<2>(a.partB, (int)blockIdx.x, 1, 1, acc);
}

__global__ __launch_bounds__(BLOCK) void k_mcount(MxArgs a) {
    __shared__ double sh[BLOCK / 64];
    double v[1] = {0.0};
    for (int s = threadIdx.x; s < a.nsys; s += BLOCK) v[0] += a.stat[1][s].done ? 1.0 : 0.0;
    block_sum<1>(v, sh);
    if (threadIdx.x == 0) *a.ndone = (int)v[0];
}

// ---------------------------------------------------------------------------
// Result: x - mean(x) on every singular component (the gauge fix).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_mproject(MxArgs a, int tile) {
    const TileWalk w = my_tiles(a.ntiles);
    for (int t = w.first; t < w.end;) {
        const TileDesc T = a.tiles[t];
        const int nrun = w.run(t, T.tile_end);
        const MxStat& S = a.stat[1][T.sys];
        const double* __restrict__ x = a.X[S.xbuf];
        double mean[MAX_COMP];
#pragma unroll
        for (int q = 0; q < MAX_COMP; ++q) mean[q] = S.mean_x[q];
        for (int u = 0; u < nrun; ++u) {
            const int base = (t + u * w.stride - T.tile_begin) * tile;
            const int c_end = min(T.n, base + tile);
            for (int c = base + (int)threadIdx.x; c < c_end; c += BLOCK) {
                const long o = T.off + c;
                const int slot = a.comp[o];
                double mu = 0.0;
#pragma unroll
                for (int q = 0; q < MAX_COMP; ++q) mu = slot == q ? mean[q] : mu;
                a.xout[o] = x[o] - mu;
            }
        }
        t += nrun * w.stride;
    }
}

constexpr long SINK_CELLS = 64;   // past the arenas of p and q: the column stencil's discarded stores

struct MixedWork {
    int fmt = -1;
    DevBuf<double> X0, X1, r64, xout, partR, partA, partB, sumR;
    DevBuf<unsigned char> w[3];
    DevBuf<unsigned char> e;
    DevBuf<float> scale, dinv, d, r, z, q, p0, p1;
    DevBuf<unsigned> fac;           // z-line factors (allocated on first use)
    DevBuf<MxDyn> dyn0, dyn1;
    DevBuf<MxStat> stat0, stat1;
    DevBuf<MxConst> cst;
    DevBuf<int> ndone;
    int* host_cnt = nullptr;        // pinned, 2 slots
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipEvent_t t0 = nullptr, t1 = nullptr;
    GraphCache<MxArgs> graphs;      // one instantiated graph per (format, interval, precond, shape, buffers)
    hipGraphExec_t exec = nullptr;  // the cached graph of the current solve
    MxArgs args{};
};

void free_mixed(MixedWork* w) {
    if (!w) return;
    w->graphs.clear();
    for (auto& e : w->ev)
        if (e) (void)hipEventDestroy(e);
    if (w->t0) (void)hipEventDestroy(w->t0);
    if (w->t1) (void)hipEventDestroy(w->t1);
    if (w->host_cnt) (void)hipHostFree(w->host_cnt);
    delete w;
}

int download_x(Batch& B, int s, const double* x_compact);

// First-time buffers of the mixed solver; the caller publishes w only when
// every one of them exists.
static int init_mixed(Batch& B, MixedWork* w) {
    {
        const size_t n = B.ncells;
        EAFP_TRY_RC(w->X0.alloc(n));
        EAFP_TRY_RC(w->X1.alloc(n));
        EAFP_TRY_RC(w->r64.alloc(n));
        EAFP_TRY_RC(w->xout.alloc(n));
        EAFP_TRY_RC(w->dinv.alloc(n));
        EAFP_TRY_RC(w->d.alloc(n));
        EAFP_TRY_RC(w->r.alloc(n));
        EAFP_TRY_RC(w->z.alloc(n));
        EAFP_TRY_RC(w->q.alloc(n + SINK_CELLS));
        EAFP_TRY_RC(w->p0.alloc(n + SINK_CELLS));
        EAFP_TRY_RC(w->p1.alloc(n + SINK_CELLS));
        EAFP_TRY(hipMemset(w->d.p, 0, n * sizeof(float)));
        EAFP_TRY(hipMemset(w->p0.p, 0, (n + SINK_CELLS) * sizeof(float)));
        EAFP_TRY(hipMemset(w->p1.p, 0, (n + SINK_CELLS) * sizeof(float)));
        EAFP_TRY_RC(w->dyn0.alloc(B.nsys));
        EAFP_TRY_RC(w->dyn1.alloc(B.nsys));
        EAFP_TRY_RC(w->stat0.alloc(B.nsys));
        EAFP_TRY_RC(w->stat1.alloc(B.nsys));
        EAFP_TRY_RC(w->cst.alloc(B.nsys));
        EAFP_TRY_RC(w->ndone.alloc(1));
        EAFP_TRY(hipHostMalloc((void**)&w->host_cnt, 2 * sizeof(int), hipHostMallocDefault));
        for (auto& e : w->ev) EAFP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        EAFP_TRY(hipEventCreate(&w->t0));
        EAFP_TRY(hipEventCreate(&w->t1));
    }
    return EAFP_OK;
}

static int ensure_mixed(Batch& B, int fmt) {
    if (!B.mixed) {
        MixedWork* nw = new MixedWork();
        const int rc = init_mixed(B, nw);
        if (rc != EAFP_OK) {
            free_mixed(nw);
            return rc;
        }
        B.mixed = nw;
    }
    MixedWork* w = B.mixed;
    if (w->fmt != fmt) {
        const size_t n = B.ncells;
        w->fmt = -1;   // a failed allocation below leaves the weights invalid, not stale
        for (auto& wb : w->w) EAFP_TRY_RC(wb.alloc(n * wbytes(fmt)));
        EAFP_TRY_RC(w->e.alloc(n * wbytes(fmt)));
        if (fmt == EAFP_COEF_FP16) EAFP_TRY_RC(w->scale.alloc(n));
        w->fmt = fmt;
    }
    EAFP_TRY_RC(w->partR.alloc((size_t)B.ntiles * NRES));
    EAFP_TRY_RC(w->sumR.alloc((size_t)B.nsys * NRES));
    EAFP_TRY_RC(w->partA.alloc((size_t)B.ntiles));
    EAFP_TRY_RC(w->partB.alloc((size_t)std::max(B.ntiles, B.nltiles) * 2));
    if (B.nltiles > 0 && !w->fac.p) EAFP_TRY_RC(w->fac.alloc(B.ncells));
    MxArgs& a = w->args;
    a.tiles = B.tiles.p;
    a.ntiles = B.ntiles;
    a.nsys = B.nsys;
    a.cols = B.cols.p;
    a.aW = B.coef[C_AW].p;
    a.aE = B.coef[C_AE].p;
    a.aS = B.coef[C_AS].p;
    a.aN = B.coef[C_AN].p;
    a.aB = B.coef[C_AB].p;
    a.aT = B.coef[C_AT].p;
    a.aP = B.coef[C_AP].p;
    a.b = B.b.p;
    a.fold = B.fold.p;
    a.x0 = B.x0.p;
    a.comp = B.comp.p;
    a.X[0] = w->X0.p;
    a.X[1] = w->X1.p;
    a.r64 = w->r64.p;
    a.xout = w->xout.p;
    for (int c = 0; c < 3; ++c) a.w[c] = w->w[c].p;
    a.scale = fmt == EAFP_COEF_FP16 ? w->scale.p : nullptr;
    a.e = w->e.p;
    a.dinv = w->dinv.p;
    a.lt = B.ltiles.p;
    a.fac = w->fac.p;
    a.state_first = B.nsys > 1 ? 1 : 0;
    a.d = w->d.p;
    a.r = w->r.p;
    a.z = w->z.p;
    a.q = w->q.p;
    a.p[0] = w->p0.p;
    a.p[1] = w->p1.p;
    a.sink = B.ncells;   // a multiple of 32 cells: 16-byte aligned
    a.partR = w->partR.p;
    a.sumR = w->sumR.p;
    a.geom = B.geom_d.p;
    a.partA = w->partA.p;
    a.partB = w->partB.p;
    a.dyn[0] = w->dyn0.p;
    a.dyn[1] = w->dyn1.p;
    a.stat[0] = w->stat0.p;
    a.stat[1] = w->stat1.p;
    a.cst = w->cst.p;
    a.ndone = w->ndone.p;
    return EAFP_OK;
}

template <int F, int CPT>
static void launch_mspmv_t(Batch& B, int par) {
    k_mspmv<F, CPT><<<B.grid, BLOCK, 0, B.stream>>>(B.mixed->args, par);
}
template <int CPT>
static void launch_mupdate_t(Batch& B, int par) {
    k_mupdate<CPT><<<B.grid, BLOCK, 0, B.stream>>>(B.mixed->args, par);
}

template <int F>
static void launch_mspmv_f(Batch& B, int par) {
    if (B.col_split > 0) {
        if (B.quad) k_mspmv_col<F><<<(int)B.cols_h.size(), BLOCK, 0, B.stream>>>(B.mixed->args, par);
        else k_mspmv_colg<F><<<(int)B.cols_h.size(), BLOCK, 0, B.stream>>>(B.mixed->args, par);
        return;
    }
    if (B.quad && B.cpt >= 4) {
        if (B.cpt == 8) k_mspmv4<F, 2><<<B.grid, BLOCK, 0, B.stream>>>(B.mixed->args, par);
        else k_mspmv4<F, 1><<<B.grid, BLOCK, 0, B.stream>>>(B.mixed->args, par);
        return;
    }
    switch (B.cpt) {
        case 1: launch_mspmv_t<F, 1>(B, par); break;
        case 4: launch_mspmv_t<F, 4>(B, par); break;
        case 8: launch_mspmv_t<F, 8>(B, par); break;
        default: launch_mspmv_t<F, 2>(B, par); break;
    }
}

static void launch_mspmv(Batch& B, int fmt, int par) {
    if (fmt == EAFP_COEF_BF16) launch_mspmv_f<EAFP_COEF_BF16>(B, par);
    else if (fmt == EAFP_COEF_FP16) launch_mspmv_f<EAFP_COEF_FP16>(B, par);
    else launch_mspmv_f<EAFP_COEF_FP32>(B, par);
}

static void launch_mupdate(Batch& B, int par) {
    if (B.mixed->args.line) {
        if (B.quad) k_mupdate_line<true><<<B.nltiles, BLOCK, 0, B.stream>>>(B.mixed->args, par);
        else k_mupdate_line<false><<<B.nltiles, BLOCK, 0, B.stream>>>(B.mixed->args, par);
        return;
    }
    if (B.quad && B.cpt >= 4) {
        if (B.cpt == 8) k_mupdate4<2><<<B.grid, BLOCK, 0, B.stream>>>(B.mixed->args, par);
        else k_mupdate4<1><<<B.grid, BLOCK, 0, B.stream>>>(B.mixed->args, par);
        return;
    }
    switch (B.cpt) {
        case 1: launch_mupdate_t<1>(B, par); break;
        case 4: launch_mupdate_t<4>(B, par); break;
        case 8: launch_mupdate_t<8>(B, par); break;
        default: launch_mupdate_t<2>(B, par); break;
    }
}

static void launch_refine(Batch& B) {
    const int tile = BLOCK * B.cpt;
    k_res<<<B.grid, BLOCK, 0, B.stream>>>(B.mixed->args, tile);
    k_rsum<<<B.nsys, BLOCK, 0, B.stream>>>(B.mixed->args);
    if (B.mixed->args.line && B.quad) k_mreplace_line<true><<<B.nltiles, BLOCK, 0, B.stream>>>(B.mixed->args);
    else if (B.mixed->args.line) k_mreplace_line<false><<<B.nltiles, BLOCK, 0, B.stream>>>(B.mixed->args);
    else k_mreplace<<<B.grid, BLOCK, 0, B.stream>>>(B.mixed->args, tile);
}

// The graph of one refinement step for every system.  Looked up in the
// workspace's cache by format, interval, preconditioner, launch shape and
// every buffer its kernels name; captured and instantiated on a miss.
static int mixed_graph(Batch& B, int fmt, int interval) {
    MixedWork* w = B.mixed;
    const int key[GraphCache<MxArgs>::kKeys] = {fmt, interval, w->args.line, 0};
    w->exec = w->graphs.find(key, B.tiling_gen, w->args);
    if (w->exec) return EAFP_OK;
    EAFP_TRY(hipStreamSynchronize(B.stream));   // no earlier launch of a graph that make_room may drop
    w->graphs.make_room(B.tiling_gen);
    hipGraph_t graph = nullptr;
    EAFP_TRY(hipStreamBeginCapture(B.stream, hipStreamCaptureModeThreadLocal));
    for (int c = 0; c < interval; ++c) {
        launch_mspmv(B, fmt, c & 1);
        launch_mupdate(B, c & 1);
    }
    launch_refine(B);
    k_mcount<<<1, BLOCK, 0, B.stream>>>(w->args);
    hipError_t e = hipStreamEndCapture(B.stream, &graph);
    if (e != hipSuccess) return (int)e;
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return (int)e;
    w->graphs.add(key, B.tiling_gen, w->args, exec);
    w->exec = exec;
    counter_add(CTR_CAPTURES, 1);
    return EAFP_OK;
}

static int mixed_start(Batch& B, const eafp_mixed_opts& o, int fmt) {
    MixedWork* w = B.mixed;
    std::vector<MxConst> cst(B.nsys);
    for (int s = 0; s < B.nsys; ++s) {
        MxConst& K = cst[s];
        K = MxConst{};
        K.tol = o.tol;
        K.max_steps = o.max_steps;
        K.restart = o.restart ? 1 : 0;
        K.ncomp = B.comp_info[s].ncomp;
        for (int q = 0; q < MAX_COMP; ++q) K.inv_cnt[q] = B.comp_info[s].inv_cnt[q];
    }
    std::vector<MxStat> st(B.nsys, MxStat{});
    std::vector<MxDyn> dy(B.nsys, MxDyn{});
    EAFP_TRY(hipMemcpyAsync(w->cst.p, cst.data(), sizeof(MxConst) * B.nsys, hipMemcpyHostToDevice, B.stream));
    EAFP_TRY(hipMemcpyAsync(w->stat1.p, st.data(), sizeof(MxStat) * B.nsys, hipMemcpyHostToDevice, B.stream));
    EAFP_TRY(hipMemcpyAsync(w->dyn1.p, dy.data(), sizeof(MxDyn) * B.nsys, hipMemcpyHostToDevice, B.stream));
    // the sources are this function's vectors: do not leave the copies queued behind them
    EAFP_TRY(hipStreamSynchronize(B.stream));
    const int tile = BLOCK * B.cpt;
    if (fmt == EAFP_COEF_BF16) k_compress<EAFP_COEF_BF16><<<B.grid, BLOCK, 0, B.stream>>>(w->args, tile);
    else if (fmt == EAFP_COEF_FP16) k_compress<EAFP_COEF_FP16><<<B.grid, BLOCK, 0, B.stream>>>(w->args, tile);
    else k_compress<EAFP_COEF_FP32><<<B.grid, BLOCK, 0, B.stream>>>(w->args, tile);
    if (w->args.line) {
        if (fmt == EAFP_COEF_BF16) k_linefac<EAFP_COEF_BF16><<<B.nltiles, BLOCK, 0, B.stream>>>(w->args);
        else if (fmt == EAFP_COEF_FP16) k_linefac<EAFP_COEF_FP16><<<B.nltiles, BLOCK, 0, B.stream>>>(w->args);
        else k_linefac<EAFP_COEF_FP32><<<B.nltiles, BLOCK, 0, B.stream>>>(w->args);
    }
    launch_refine(B);    // x = x0, first residual, first inner start
    EAFP_TRY(hipGetLastError());
    return EAFP_OK;
}

// Argument and state checks, then the singular components of every system
// uploaded since its last labelling.
static int check_mixed(Batch& B, const eafp_mixed_opts& o) {
    for (int s = 0; s < B.nsys; ++s)
        if (!B.uploaded[s]) return EAFP_ERR_STATE;
    if (!(o.tol > 0.0) || o.max_steps < 1 || o.interval < 0) return EAFP_ERR_ARG;
    if (o.coef_format < EAFP_COEF_FP32 || o.coef_format > EAFP_COEF_FP16) return EAFP_ERR_ARG;
    if (o.precond != EAFP_PRECOND_JACOBI && o.precond != EAFP_PRECOND_ZLINE) return EAFP_ERR_ARG;
    if (o.precond == EAFP_PRECOND_ZLINE && B.nltiles == 0) return EAFP_ERR_UNSUPPORTED;
    EAFP_TRY_RC(B.label_components());
    for (int s = 0; s < B.nsys; ++s)
        if (B.comp_info[s].overflow) return EAFP_ERR_COMPONENTS;
    return EAFP_OK;
}

int solve_mixed(Batch& B, const eafp_mixed_opts& o, eafp_stats* stats) {
    EAFP_TRY_RC(check_mixed(B, o));
    const int fmt = o.coef_format;
    EAFP_TRY_RC(ensure_mixed(B, fmt));
    MixedWork* w = B.mixed;
    w->args.line = o.precond == EAFP_PRECOND_ZLINE ? 1 : 0;
    int interval = o.interval > 0 ? o.interval : 32;
    interval = (interval + 1) & ~1;    // even: the p buffer parity restarts every step
    EAFP_TRY_RC(mixed_graph(B, fmt, interval));
    EAFP_TRY(hipEventRecord(w->t0, B.stream));
    EAFP_TRY_RC(mixed_start(B, o, fmt));

    // One graph launch = one refinement step for every system; poll the
    // finished count of the step before the one just launched.
    const long max_launch = (long)o.max_steps + 2;
    long launched = 0;
    for (;;) {
        if (launched < max_launch) {
            EAFP_TRY(hipGraphLaunch(w->exec, B.stream));
            counter_add(CTR_LAUNCHES, 1);
            const int slot = (int)(launched & 1);
            EAFP_TRY(hipMemcpyAsync(&w->host_cnt[slot], w->ndone.p, sizeof(int), hipMemcpyDeviceToHost, B.stream));
            EAFP_TRY(hipEventRecord(w->ev[slot], B.stream));
            ++launched;
        }
        if (launched >= 2) {
            const int older = (int)(launched & 1);
            EAFP_TRY(hipEventSynchronize(w->ev[older]));
            if (w->host_cnt[older] >= B.nsys) break;
        }
        if (launched >= max_launch) {
            EAFP_TRY(hipEventSynchronize(w->ev[(launched - 1) & 1]));
            break;
        }
    }
    k_mproject<<<B.grid, BLOCK, 0, B.stream>>>(w->args, BLOCK * B.cpt);
    EAFP_TRY(hipGetLastError());
    for (int s = 0; s < B.nsys; ++s) EAFP_TRY_RC(download_x(B, s, w->xout.p));
    EAFP_TRY(hipEventRecord(w->t1, B.stream));
    std::vector<MxStat> st(B.nsys);
    std::vector<MxDyn> dy(B.nsys);
    EAFP_TRY(hipMemcpyAsync(st.data(), w->stat1.p, sizeof(MxStat) * B.nsys, hipMemcpyDeviceToHost, B.stream));
    EAFP_TRY(hipMemcpyAsync(dy.data(), w->dyn1.p, sizeof(MxDyn) * B.nsys, hipMemcpyDeviceToHost, B.stream));
    EAFP_TRY(hipStreamSynchronize(B.stream));
    float ms = 0.f;
    EAFP_TRY(hipEventElapsedTime(&ms, w->t0, w->t1));
    if (stats) {
        for (int s = 0; s < B.nsys; ++s) {
            eafp_stats& out = stats[s];
            out = eafp_stats{};
            out.residual = st[s].res_l2;
            out.n_iter = dy[s].total_inner;
            out.n_refine = st[s].n_refine;
            out.status = st[s].done ? st[s].status : EAFP_STOP_RUNNING;
            out.true_rel_l2 = st[s].res_l2;
            out.true_rel_l1 = st[s].res_l1;
            out.seconds = ms * 1e-3;
        }
    }
    return EAFP_OK;
}

// Average durations of the two inner-iteration kernels over `reps`
// iterations (systems kept running: the tolerance is out of reach), each
// launch bracketed by HIP events on the library stream.
int time_mixed_kernels(Batch& B, int fmt, int precond, int reps, float* ms_spmv, float* ms_update) {
    eafp_mixed_opts o{};
    o.tol = 1e-300;
    o.interval = 2;
    o.max_steps = 1 << 20;
    o.coef_format = fmt;
    o.precond = precond;
    EAFP_TRY_RC(check_mixed(B, o));
    if (reps < 1) return EAFP_ERR_ARG;
    EAFP_TRY_RC(ensure_mixed(B, fmt));
    B.mixed->args.line = precond == EAFP_PRECOND_ZLINE ? 1 : 0;
    EAFP_TRY_RC(mixed_start(B, o, fmt));
    hipEvent_t e[3];
    for (auto& x : e) EAFP_TRY(hipEventCreate(&x));
    float tot_s = 0.f, tot_u = 0.f;
    for (int r = 0; r < reps; ++r) {
        const int par = r & 1;
        EAFP_TRY(hipEventRecord(e[0], B.stream));
        launch_mspmv(B, fmt, par);
        EAFP_TRY(hipEventRecord(e[1], B.stream));
        launch_mupdate(B, par);
        EAFP_TRY(hipEventRecord(e[2], B.stream));
        EAFP_TRY(hipEventSynchronize(e[2]));
        float ta = 0.f, tb = 0.f;
        EAFP_TRY(hipEventElapsedTime(&ta, e[0], e[1]));
        EAFP_TRY(hipEventElapsedTime(&tb, e[1], e[2]));
        tot_s += ta;
        tot_u += tb;
    }
    for (auto& x : e) (void)hipEventDestroy(x);
    if (ms_spmv) *ms_spmv = tot_s / (float)reps;
    if (ms_update) *ms_update = tot_u / (float)reps;
    return EAFP_OK;
}

}  // namespace eafp
