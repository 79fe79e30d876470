// cg64.hip -- FP64 Jacobi-preconditioned CG, the algorithm of cg_3d_mpi
// (src/mod_solver_3d.f90:485-620) on the GPU, for one or many systems.
//
// Per iteration two kernels, each one streaming pass over the cells:
//
//   k64_spmv    pv_new = z + beta*pv_old (recomputed at the 6 neighbours, so
//               no separate xpay pass), q = A*pv_new, partial p.q
//               finisher: |p.q| < SMALL -> breakdown, else alpha = rz/p.q
//   k64_update  phi += alpha*pv_new, r -= alpha*q, z = r/max(|aP|,SMALL),
//               partials r.r and r.z
//               finisher: residual = ||r||/||Su||, n_iter = it, stop test,
//               beta = rz_new/max(rz,SMALL)
//
// The scalar recurrences follow the reference line by line (same tests in
// the same order, including the NaN behaviour: a NaN residual never passes
// `residual < tol`).  pv is double-buffered because neighbours read the old
// value while the owner writes the new one.  Iterations are replayed from
// a captured hipGraph in chunks; every kernel of a finished system returns
// at once, and the host polls the number of finished systems between
// chunks.
//
// Algorithmic HBM bytes per cell and iteration (FP64): k64_spmv reads 7
// coefficients + z + pv_old and writes pv_new + q (88 B); k64_update reads
// phi, r, pv_new, q, aP and writes phi, r, z (64 B): 152 B/cell/iter.
#include <chrono>
#include <vector>

#include "batch.hpp"

namespace eafp {

struct CgState {
    double bnorm, residual, rz, alpha, beta, tol;
    int n_iter, max_iter, done, status;
    unsigned counter;
    unsigned pad;
};

struct Cg64Args {
    const SysGeom* geom;
    const int* tile_sys;
    const int* tile_start;
    int ntiles;
    const double *aW, *aE, *aS, *aN, *aB, *aT, *aP;
    const double *b, *x0;
    double *x, *r, *z, *q;
    double* p[2];
    double* partial;     // ntiles * 3
    CgState* st;
    int* ndone;
};

struct Cg64Work {
    DevBuf<double> x, r, z, q, p0, p1, partial;
    DevBuf<CgState> st;
    DevBuf<int> ndone;
    int* ndone_host = nullptr;     // pinned, 2 slots
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipEvent_t t0 = nullptr, t1 = nullptr;
    hipGraphExec_t exec = nullptr;
    int exec_chunk = 0;
    Cg64Args args;
};

void free_cg64(Cg64Work* w) {
    if (!w) return;
    if (w->exec) (void)hipGraphExecDestroy(w->exec);
    for (auto& e : w->ev)
        if (e) (void)hipEventDestroy(e);
    if (w->t0) (void)hipEventDestroy(w->t0);
    if (w->t1) (void)hipEventDestroy(w->t1);
    if (w->ndone_host) (void)hipHostFree(w->ndone_host);
    delete w;
}

int download_x(Batch& B, int s, const double* x_compact);

// Value of the stencil neighbour at compact offset o + d, or 0 outside
// the r/z walls (the zeroed halo of the reference's work vectors).
#define EAFP_NEIGHBOURS(id, g, o)                                           \
    const long sr_ = (long)(g).nth * (g).nz;                                \
    const long oW = (o) - sr_, oE = (o) + sr_;                              \
    const long oS = (id).j > 0 ? (o) - (g).nz : (o) + (long)((g).nth - 1) * (g).nz; \
    const long oN = (id).j < (g).nth - 1 ? (o) + (g).nz : (o) - (long)((g).nth - 1) * (g).nz; \
    const long oB = (o) - 1, oT = (o) + 1;                                  \
    const bool hW = (id).i > 0, hE = (id).i < (g).nr - 1;                   \
    const bool hB = (id).k > 0, hT = (id).k < (g).nz - 1;

// ---------------------------------------------------------------------------
// r = Su - A*phi0, z = r/max(|aP|,SMALL), partials Su.Su, r.r, r.z
// (src/mod_solver_3d.f90:509-530).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k64_init(Cg64Args a) {
    __shared__ double sh[3 * (BLOCK / 64)];
    __shared__ int flag;
    for (TileWalk w(a.ntiles); w.valid(); w.next()) {
        const int s = a.tile_sys[w.t];
        const SysGeom g = a.geom[s];
        const int start = a.tile_start[w.t];
        double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int m = 0; m < CELLS_PER_THREAD; ++m) {
            const int c = start + m * BLOCK + (int)threadIdx.x;
            if (c < g.n) {
                const CellIdx id = decompose(c, g);
                const long o = g.off + c;
                EAFP_NEIGHBOURS(id, g, o)
                const double* x = a.x0;
                const double xW = hW ? x[oW] : 0.0, xE = hE ? x[oE] : 0.0;
                const double xB = hB ? x[oB] : 0.0, xT = hT ? x[oT] : 0.0;
                const double ap = a.aP[o];
                const double ax = ap * x[o] - a.aW[o] * xW - a.aE[o] * xE - a.aS[o] * x[oS] -
                                  a.aN[o] * x[oN] - a.aB[o] * xB - a.aT[o] * xT;
                const double su = a.b[o];
                const double rr = su - ax;
                const double zz = rr / fmax(fabs(ap), SMALL);
                a.x[o] = x[o];
                a.r[o] = rr;
                a.z[o] = zz;
                acc[0] += su * su;
                acc[1] += rr * rr;
                acc[2] += rr * zz;
            }
        }
        block_sum<3>(acc, sh);
        if (threadIdx.x == 0) {
#pragma unroll
            for (int c = 0; c < 3; ++c) a.partial[(long)w.t * 3 + c] = acc[c];
        }
        CgState& S = a.st[s];
        if (arrive_last(&S.counter, (unsigned)(g.tile_end - g.tile_begin), &flag)) {
            double tot[3];
            reduce_tiles<3>(a.partial, 3, g.tile_begin, g.tile_end, tot, sh);
            if (threadIdx.x == 0) {
                S.counter = 0;
                const double bnorm = sqrt(tot[0]);
                S.bnorm = bnorm;
                S.n_iter = 0;
                S.beta = 0.0;
                if (bnorm < SMALL) {
                    S.residual = 0.0;
                    S.done = 1;
                    S.status = EAFP_STOP_ZERO_RHS;
                    atomicAdd(a.ndone, 1);
                } else {
                    S.rz = tot[2];
                    S.residual = sqrt(tot[1]) / bnorm;
                    if (S.max_iter <= 0) {
                        S.done = 1;
                        S.status = EAFP_STOP_MAX_ITER;
                        atomicAdd(a.ndone, 1);
                    }
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// pv = z + beta*pv ; q = A*pv ; p.q  (src/mod_solver_3d.f90:533-537,556-558)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k64_spmv(Cg64Args a) {
    __shared__ double sh[1 * (BLOCK / 64)];
    __shared__ int flag;
    for (TileWalk w(a.ntiles); w.valid(); w.next()) {
        const int s = a.tile_sys[w.t];
        CgState& S = a.st[s];
        if (S.done) continue;
        const SysGeom g = a.geom[s];
        const int start = a.tile_start[w.t];
        const int it = S.n_iter + 1;
        const double beta = S.beta;
        const bool first = (it == 1);
        const double* __restrict__ pold = a.p[(it - 1) & 1];
        double* __restrict__ pnew = a.p[it & 1];
        const double* __restrict__ z = a.z;
        double acc[1] = {0.0};
#pragma unroll
        for (int m = 0; m < CELLS_PER_THREAD; ++m) {
            const int c = start + m * BLOCK + (int)threadIdx.x;
            if (c < g.n) {
                const CellIdx id = decompose(c, g);
                const long o = g.off + c;
                EAFP_NEIGHBOURS(id, g, o)
                auto pv = [&](long q) { return first ? z[q] : z[q] + beta * pold[q]; };
                const double pc = pv(o);
                const double pW = hW ? pv(oW) : 0.0, pE = hE ? pv(oE) : 0.0;
                const double pB = hB ? pv(oB) : 0.0, pT = hT ? pv(oT) : 0.0;
                const double pS = pv(oS), pN = pv(oN);
                const double y = a.aP[o] * pc - a.aW[o] * pW - a.aE[o] * pE - a.aS[o] * pS -
                                 a.aN[o] * pN - a.aB[o] * pB - a.aT[o] * pT;
                pnew[o] = pc;
                a.q[o] = y;
                acc[0] += pc * y;
            }
        }
        block_sum<1>(acc, sh);
        if (threadIdx.x == 0) a.partial[(long)w.t * 3] = acc[0];
        if (arrive_last(&S.counter, (unsigned)(g.tile_end - g.tile_begin), &flag)) {
            double tot[1];
            reduce_tiles<1>(a.partial, 3, g.tile_begin, g.tile_end, tot, sh);
            if (threadIdx.x == 0) {
                S.counter = 0;
                const double pq = tot[0];
                if (fabs(pq) < SMALL) {
                    S.done = 1;
                    S.status = EAFP_STOP_BREAKDOWN;
                    atomicAdd(a.ndone, 1);
                } else {
                    S.alpha = S.rz / pq;
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// phi += alpha*pv ; r -= alpha*q ; residual ; z = r/max(|aP|,SMALL) ; r.z ;
// beta  (src/mod_solver_3d.f90:538-555)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k64_update(Cg64Args a) {
    __shared__ double sh[2 * (BLOCK / 64)];
    __shared__ int flag;
    for (TileWalk w(a.ntiles); w.valid(); w.next()) {
        const int s = a.tile_sys[w.t];
        CgState& S = a.st[s];
        if (S.done) continue;
        const SysGeom g = a.geom[s];
        const int start = a.tile_start[w.t];
        const int it = S.n_iter + 1;
        const double alpha = S.alpha;
        const double* __restrict__ pv = a.p[it & 1];
        double acc[2] = {0.0, 0.0};
#pragma unroll
        for (int m = 0; m < CELLS_PER_THREAD; ++m) {
            const int c = start + m * BLOCK + (int)threadIdx.x;
            if (c < g.n) {
                const long o = g.off + c;
                a.x[o] = a.x[o] + alpha * pv[o];
                const double rr = a.r[o] - alpha * a.q[o];
                const double zz = rr / fmax(fabs(a.aP[o]), SMALL);
                a.r[o] = rr;
                a.z[o] = zz;
                acc[0] += rr * rr;
                acc[1] += rr * zz;
            }
        }
        block_sum<2>(acc, sh);
        if (threadIdx.x == 0) {
            a.partial[(long)w.t * 3 + 1] = acc[0];
            a.partial[(long)w.t * 3 + 2] = acc[1];
        }
        if (arrive_last(&S.counter, (unsigned)(g.tile_end - g.tile_begin), &flag)) {
            double tot[2];
            // partial column 1..2 of every tile
            tot[0] = 0.0;
            tot[1] = 0.0;
            for (int t = g.tile_begin + (int)threadIdx.x; t < g.tile_end; t += BLOCK) {
                tot[0] += a.partial[(long)t * 3 + 1];
                tot[1] += a.partial[(long)t * 3 + 2];
            }
            block_sum<2>(tot, sh);
            if (threadIdx.x == 0) {
                S.counter = 0;
                const double res = sqrt(tot[0]) / S.bnorm;
                S.residual = res;
                S.n_iter = it;
                if (res < S.tol) {
                    S.done = 1;
                    S.status = EAFP_STOP_CONVERGED;
                    atomicAdd(a.ndone, 1);
                } else {
                    S.beta = tot[1] / fmax(S.rz, SMALL);
                    S.rz = tot[1];
                    if (it >= S.max_iter) {
                        S.done = 1;
                        S.status = EAFP_STOP_MAX_ITER;
                        atomicAdd(a.ndone, 1);
                    }
                }
            }
        }
    }
}

static int ensure_cg64(Batch& B) {
    if (B.cg64) return EAFP_OK;
    Cg64Work* w = new Cg64Work();
    B.cg64 = w;
    const size_t n = B.ncells;
    EAFP_TRY_RC(w->x.alloc(n));
    EAFP_TRY_RC(w->r.alloc(n));
    EAFP_TRY_RC(w->z.alloc(n));
    EAFP_TRY_RC(w->q.alloc(n));
    EAFP_TRY_RC(w->p0.alloc(n));
    EAFP_TRY_RC(w->p1.alloc(n));
    EAFP_TRY_RC(w->partial.alloc((size_t)B.ntiles * 3));
    EAFP_TRY_RC(w->st.alloc(B.nsys));
    EAFP_TRY_RC(w->ndone.alloc(1));
    EAFP_TRY(hipHostMalloc((void**)&w->ndone_host, 2 * sizeof(int), hipHostMallocDefault));
    for (auto& e : w->ev) EAFP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    EAFP_TRY(hipEventCreate(&w->t0));
    EAFP_TRY(hipEventCreate(&w->t1));
    Cg64Args& a = w->args;
    a.geom = B.geom_d.p;
    a.tile_sys = B.tile_sys.p;
    a.tile_start = B.tile_start.p;
    a.ntiles = B.ntiles;
    a.aW = B.coef[C_AW].p;
    a.aE = B.coef[C_AE].p;
    a.aS = B.coef[C_AS].p;
    a.aN = B.coef[C_AN].p;
    a.aB = B.coef[C_AB].p;
    a.aT = B.coef[C_AT].p;
    a.aP = B.coef[C_AP].p;
    a.b = B.b.p;
    a.x0 = B.x0.p;
    a.x = w->x.p;
    a.r = w->r.p;
    a.z = w->z.p;
    a.q = w->q.p;
    a.p[0] = w->p0.p;
    a.p[1] = w->p1.p;
    a.partial = w->partial.p;
    a.st = w->st.p;
    a.ndone = w->ndone.p;
    // zero halos of nothing: compact vectors have no halos, but keep the
    // unused tail cells of each system (alignment padding) at zero
    EAFP_TRY(hipMemset(w->p0.p, 0, n * sizeof(double)));
    EAFP_TRY(hipMemset(w->p1.p, 0, n * sizeof(double)));
    return EAFP_OK;
}

static int cg64_graph(Batch& B, int chunk) {
    Cg64Work* w = B.cg64;
    if (w->exec && w->exec_chunk == chunk) return EAFP_OK;
    if (w->exec) {
        (void)hipGraphExecDestroy(w->exec);
        w->exec = nullptr;
    }
    hipGraph_t graph = nullptr;
    EAFP_TRY(hipStreamBeginCapture(B.stream, hipStreamCaptureModeThreadLocal));
    for (int c = 0; c < chunk; ++c) {
        k64_spmv<<<B.grid, BLOCK, 0, B.stream>>>(w->args);
        k64_update<<<B.grid, BLOCK, 0, B.stream>>>(w->args);
    }
    hipError_t e = hipStreamEndCapture(B.stream, &graph);
    if (e != hipSuccess) return (int)e;
    e = hipGraphInstantiate(&w->exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return (int)e;
    w->exec_chunk = chunk;
    return EAFP_OK;
}

int solve_fp64(Batch& B, int max_iter, double tol, eafp_stats* stats, int chunk) {
    for (int s = 0; s < B.nsys; ++s)
        if (!B.uploaded[s]) return EAFP_ERR_STATE;
    EAFP_TRY_RC(ensure_cg64(B));
    Cg64Work* w = B.cg64;
    if (chunk <= 0) chunk = 16;
    EAFP_TRY_RC(cg64_graph(B, chunk));

    std::vector<CgState> st(B.nsys);
    for (auto& S : st) {
        S = CgState{};
        S.tol = tol;
        S.max_iter = max_iter;
    }
    EAFP_TRY(hipMemcpyAsync(w->st.p, st.data(), sizeof(CgState) * B.nsys, hipMemcpyHostToDevice, B.stream));
    EAFP_TRY(hipMemsetAsync(w->ndone.p, 0, sizeof(int), B.stream));
    EAFP_TRY(hipEventRecord(w->t0, B.stream));
    k64_init<<<B.grid, BLOCK, 0, B.stream>>>(w->args);
    EAFP_TRY(hipGetLastError());

    // Chunks of `chunk` iterations; poll the finished count of the chunk
    // before the one just launched so the queue never drains.
    const long max_chunks = (long)(max_iter > 0 ? max_iter : 0) / chunk + 2;
    long launched = 0;
    for (;;) {
        if (launched < max_chunks) {
            EAFP_TRY(hipGraphLaunch(w->exec, B.stream));
            const int slot = (int)(launched & 1);
            EAFP_TRY(hipMemcpyAsync(&w->ndone_host[slot], w->ndone.p, sizeof(int), hipMemcpyDeviceToHost, B.stream));
            EAFP_TRY(hipEventRecord(w->ev[slot], B.stream));
            ++launched;
        }
        const int prev = (int)((launched - 1) & 1);
        const int older = (int)(launched & 1);
        if (launched >= 2) {
            EAFP_TRY(hipEventSynchronize(w->ev[older]));
            if (w->ndone_host[older] >= B.nsys) break;
        }
        if (launched >= max_chunks) {
            EAFP_TRY(hipEventSynchronize(w->ev[prev]));
            break;
        }
    }
    for (int s = 0; s < B.nsys; ++s) EAFP_TRY_RC(download_x(B, s, w->x.p));
    EAFP_TRY(hipEventRecord(w->t1, B.stream));
    EAFP_TRY(hipMemcpyAsync(st.data(), w->st.p, sizeof(CgState) * B.nsys, hipMemcpyDeviceToHost, B.stream));
    EAFP_TRY(hipStreamSynchronize(B.stream));
    float ms = 0.f;
    EAFP_TRY(hipEventElapsedTime(&ms, w->t0, w->t1));
    if (stats) {
        for (int s = 0; s < B.nsys; ++s) {
            eafp_stats& o = stats[s];
            o = eafp_stats{};
            o.residual = st[s].residual;
            o.n_iter = st[s].n_iter;
            o.n_refine = 0;
            o.status = st[s].status;
            o.true_rel_l2 = -1.0;
            o.true_rel_l1 = -1.0;
            o.seconds = ms * 1e-3;
        }
    }
    return EAFP_OK;
}

}  // namespace eafp
