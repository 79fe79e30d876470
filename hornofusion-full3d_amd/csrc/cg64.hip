// cg64.hip -- FP64 Jacobi-preconditioned CG, the algorithm of cg_3d_mpi
// (src/mod_solver_3d.f90:485-620) on the GPU, for one or many systems.
//
// Per iteration two kernels, each one streaming pass over the cells:
//
//   k64_spmv    closes the previous iteration from the partial sums that
//               k64_update left (residual, stop tests, beta), then
//               pv_new = z + beta*pv_old (recomputed at the 6 neighbours,
//               so no separate xpay pass), q = A*pv_new, partials p.q
//   k64_update  from the p.q partials: breakdown test, alpha; then
//               phi += alpha*pv_new, r -= alpha*q, z = r/max(|aP|,SMALL),
//               partials r.r and r.z
//
// Every block reduces the partials it needs itself (same tile order, same
// bits in every block), so no block waits on another inside a kernel.  The
// scalar recurrences follow the reference line by line (same tests in the
// same order, including the NaN behaviour: a NaN residual never passes
// `residual < tol`).  pv is double-buffered because neighbours read the old
// value while the owner writes the new one; the buffer parity is a launch
// argument (chunks hold an even number of iterations and every system
// starts iterating at the same launch).  Iterations are replayed from a
// captured hipGraph in chunks; every kernel of a finished system skips its
// tiles, and a one-block kernel at the end of each chunk counts finished
// systems for the host.
//
// Algorithmic HBM bytes per cell and iteration (FP64): k64_spmv reads 7
// coefficients + z + pv_old and writes pv_new + q (88 B); k64_update reads
// phi, r, pv_new, q, aP and writes phi, r, z (64 B): 152 B/cell/iter.
#include <vector>

#include "batch.hpp"

namespace eafp {

struct CgSlot {
    double bnorm;      // ||Su||; < 0 until the first k64_spmv after k64_init
    double residual;   // relative recurrence residual of the last completed iteration
    double rz, alpha, beta;
    int it;            // iteration in progress (set by k64_spmv)
    int n_iter;        // completed iterations
    int done, status;
};

struct CgConst {
    double tol;
    int max_iter;
    int pad;
};

struct Cg64Args {
    const TileDesc* tiles;
    int ntiles;
    int nsys;
    const double *aW, *aE, *aS, *aN, *aB, *aT, *aP;
    const double *b, *fold, *x0;
    double *x, *r, *z, *q;
    double* p[2];
    double* partA;       // ntiles:     p.q
    double* partB;       // ntiles * 3: r.r, r.z, Su.Su
    CgSlot* slot[2];
    const CgConst* cst;
    int* ndone;          // finished systems, counted after each chunk
    int state_first;     // 1 (batches): read a system's state before its cell loads; 0: overlap them
};

struct Cg64Work {
    DevBuf<double> x, r, z, q, p0, p1, partA, partB;
    DevBuf<CgSlot> slot0, slot1;
    DevBuf<CgConst> cst;
    DevBuf<int> ndone;
    int* ndone_host = nullptr;     // pinned, 2 slots
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipEvent_t t0 = nullptr, t1 = nullptr;
    GraphCache<Cg64Args> graphs;    // one instantiated graph per (chunk, shape, buffers)
    hipGraphExec_t exec = nullptr;  // the cached graph of the current solve
    Cg64Args args{};
};

void free_cg64(Cg64Work* w) {
    if (!w) return;
    w->graphs.clear();
    for (auto& e : w->ev)
        if (e) (void)hipEventDestroy(e);
    if (w->t0) (void)hipEventDestroy(w->t0);
    if (w->t1) (void)hipEventDestroy(w->t1);
    if (w->ndone_host) (void)hipHostFree(w->ndone_host);
    delete w;
}

int download_x(Batch& B, int s, const double* x_compact);

// ---------------------------------------------------------------------------
// r = Su - A*phi0, z = r/max(|aP|,SMALL), pv_old = 0, partials r.r, r.z,
// Su.Su (src/mod_solver_3d.f90:509-530).  Writes the fresh slot 1.
// ---------------------------------------------------------------------------
template <int CPT>
__global__ __launch_bounds__(BLOCK) void k64_init(Cg64Args a) {
    constexpr int TILE = BLOCK * CPT;
    __shared__ double sh[3 * (BLOCK / 64)];
    const TileWalk w = my_tiles(a.ntiles);
    for (int t = w.first; t < w.end;) {
        const TileDesc T = a.tiles[t];
        const int nrun = w.run(t, T.tile_end);
        if (t == T.tile_begin && threadIdx.x == 0) {
            CgSlot n{};
            n.bnorm = -1.0;
            store_state(&a.slot[1][T.sys], n);
        }
        double acc[3] = {0.0, 0.0, 0.0};
        for (int u = 0; u < nrun; ++u) {
            const int base = (t + u * w.stride - T.tile_begin) * TILE;
            const int c_end = min(T.n, base + TILE);
            for (int c = base + (int)threadIdx.x; c < c_end; c += BLOCK) {
                const CellIdx id = decompose(c, T);
                const long o = T.off + c;
                const Nb nb = neighbours(id, T, o);
                const double* x = a.x0;
                const double xc = x[o], xW = x[nb.W], xE = x[nb.E], xS = x[nb.S], xN = x[nb.N], xB = x[nb.B],
                             xT = x[nb.T];
                const double ap = a.aP[o];
                const double ax = ap * xc - a.aW[o] * (nb.hW ? xW : 0.0) - a.aE[o] * (nb.hE ? xE : 0.0) -
                                  a.aS[o] * xS - a.aN[o] * xN - a.aB[o] * (nb.hB ? xB : 0.0) -
                                  a.aT[o] * (nb.hT ? xT : 0.0);
                const double su = a.b[o];
                const double rr = su - (ax - a.fold[o]);
                const double zz = rr / fmax(fabs(ap), SMALL);
                a.x[o] = xc;
                a.r[o] = rr;
                a.z[o] = zz;
                a.p[0][o] = 0.0;
                acc[0] += rr * rr;
                acc[1] += rr * zz;
                acc[2] += su * su;
            }
        }
        block_sum<3>(acc, sh);
        store_run<3>(a.partB, t, nrun, w.stride, acc);
        t += nrun * w.stride;
    }
}

// ---------------------------------------------------------------------------
// Role 0 (reads slot 1, writes slot 0).  Closes the previous iteration --
// residual, stop tests, beta (src/mod_solver_3d.f90:545-555) -- then
// pv = z + beta*pv ; q = A*pv ; p.q  (:533-534, 556-558).
// ---------------------------------------------------------------------------
template <int CPT>
__global__ __launch_bounds__(BLOCK) void k64_spmv(Cg64Args a, int par) {
    constexpr int TILE = BLOCK * CPT;
    __shared__ double sh[3 * (BLOCK / 64)];
    const double* __restrict__ pold = a.p[par];
    double* __restrict__ pnew = a.p[par ^ 1];
    const double* __restrict__ z = a.z;
    const double* __restrict__ cP = a.aP;
    const double* __restrict__ cW = a.aW;
    const double* __restrict__ cE = a.aE;
    const double* __restrict__ cS = a.aS;
    const double* __restrict__ cN = a.aN;
    const double* __restrict__ cB = a.aB;
    const double* __restrict__ cT = a.aT;
    double* __restrict__ qo = a.q;
    const TileWalk w = my_tiles(a.ntiles);
    for (int t = w.first; t < w.end;) {
        const TileDesc T = a.tiles[t];
        const int s = T.sys;
        const int nrun = w.run(t, T.tile_end);
        double zv[CPT][7], pv[CPT][7], cf[CPT][7];
        long oc[CPT];
        bool ok[CPT], mW[CPT], mE[CPT], mB[CPT], mT[CPT];
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
                mE[m] = nb.hE;
                mB[m] = nb.hB;
                mT[m] = nb.hT;
                const long idx[7] = {o, nb.W, nb.E, nb.S, nb.N, nb.B, nb.T};
#pragma unroll
                for (int q = 0; q < 7; ++q) {
                    zv[m][q] = z[idx[q]];
                    pv[m][q] = pold[idx[q]];
                }
                cf[m][0] = cP[o];
                cf[m][1] = cW[o];
                cf[m][2] = cE[o];
                cf[m][3] = cS[o];
                cf[m][4] = cN[o];
                cf[m][5] = cB[o];
                cf[m][6] = cT[o];
            }
        };
        // one system: the first batch of cell loads is in flight while the
        // partials are reduced; a batch: converged systems load nothing
        CgSlot P;
        if (a.state_first) {
            P = a.slot[1][s];
            if (!P.done) load(t);
        } else {
            load(t);
            P = a.slot[1][s];
        }
        CgSlot N = P;
        if (!P.done) {
            const CgConst K = a.cst[s];
            double tot[3];
            reduce_tiles<3>(a.partB, T.tile_begin, T.tile_end, tot, sh);
            if (P.bnorm < 0.0) {           // first iteration after k64_init
                N.bnorm = sqrt(tot[2]);
                N.n_iter = 0;
                N.beta = 0.0;
                if (N.bnorm < SMALL) {
                    N.residual = 0.0;
                    N.done = 1;
                    N.status = EAFP_STOP_ZERO_RHS;
                } else {
                    N.residual = sqrt(tot[0]) / N.bnorm;
                    N.rz = tot[1];
                    if (K.max_iter <= 0) {
                        N.done = 1;
                        N.status = EAFP_STOP_MAX_ITER;
                    }
                }
            } else {                       // iteration P.it has just been updated
                N.residual = sqrt(tot[0]) / P.bnorm;
                N.n_iter = P.it;
                if (N.residual < K.tol) {
                    N.done = 1;
                    N.status = EAFP_STOP_CONVERGED;
                } else if (P.it >= K.max_iter) {
                    N.done = 1;
                    N.status = EAFP_STOP_MAX_ITER;
                } else {
                    N.beta = tot[1] / fmax(P.rz, SMALL);
                    N.rz = tot[1];
                }
            }
            N.it = P.it + 1;
        }
        if (t == T.tile_begin && threadIdx.x == 0) store_state(&a.slot[0][s], N);
        if (!N.done) {
            const double beta = N.beta;
            double acc[1] = {0.0};
            for (int u = 0; u < nrun; ++u) {
                if (u) load(t + u * w.stride);
#pragma unroll
                for (int m = 0; m < CPT; ++m) {
                    double pp[7];
#pragma unroll
                    for (int q = 0; q < 7; ++q) pp[q] = zv[m][q] + beta * pv[m][q];
                    const double pW = mW[m] ? pp[1] : 0.0, pE = mE[m] ? pp[2] : 0.0;
                    const double pB = mB[m] ? pp[5] : 0.0, pT = mT[m] ? pp[6] : 0.0;
                    const double y = cf[m][0] * pp[0] - cf[m][1] * pW - cf[m][2] * pE - cf[m][3] * pp[3] -
                                     cf[m][4] * pp[4] - cf[m][5] * pB - cf[m][6] * pT;
                    if (ok[m]) {
                        pnew[oc[m]] = pp[0];
                        qo[oc[m]] = y;
                        acc[0] += pp[0] * y;
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
// Role 1 (reads slot 0, writes slot 1).  p.q, breakdown exit, alpha
// (src/mod_solver_3d.f90:534-536), then phi += alpha*pv ; r -= alpha*q ;
// z = r/max(|aP|,SMALL) ; partials r.r, r.z  (:538-553).
// ---------------------------------------------------------------------------
template <int CPT>
__global__ __launch_bounds__(BLOCK) void k64_update(Cg64Args a, int par) {
    constexpr int TILE = BLOCK * CPT;
    __shared__ double sh[3 * (BLOCK / 64)];
    const double* __restrict__ pv = a.p[par ^ 1];
    double* __restrict__ x = a.x;
    double* __restrict__ r = a.r;
    double* __restrict__ zo = a.z;
    const double* __restrict__ q = a.q;
    const double* __restrict__ aP = a.aP;
    const TileWalk w = my_tiles(a.ntiles);
    for (int t = w.first; t < w.end;) {
        const TileDesc T = a.tiles[t];
        const int s = T.sys;
        const int nrun = w.run(t, T.tile_end);
        double xv[CPT], pvv[CPT], rv[CPT], qv[CPT], apv[CPT];
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
                xv[m] = x[o];
                pvv[m] = pv[o];
                rv[m] = r[o];
                qv[m] = q[o];
                apv[m] = aP[o];
            }
        };
        // one system: the first batch of cell loads is in flight while the
        // partials are reduced; a batch: converged systems load nothing
        CgSlot P;
        if (a.state_first) {
            P = a.slot[0][s];
            if (!P.done) load(t);
        } else {
            load(t);
            P = a.slot[0][s];
        }
        CgSlot N = P;
        if (!P.done) {
            double tot[1];
            reduce_tiles<1>(a.partA, T.tile_begin, T.tile_end, tot, sh);
            const double pq = tot[0];
            if (fabs(pq) < SMALL) {
                N.done = 1;
                N.status = EAFP_STOP_BREAKDOWN;
            } else {
                N.alpha = P.rz / pq;
            }
        }
        if (t == T.tile_begin && threadIdx.x == 0) store_state(&a.slot[1][s], N);
        if (!N.done) {
            const double alpha = N.alpha;
            double acc[3] = {0.0, 0.0, 0.0};
            for (int u = 0; u < nrun; ++u) {
                if (u) load(t + u * w.stride);
#pragma unroll
                for (int m = 0; m < CPT; ++m) {
                    const double rr = rv[m] - alpha * qv[m];
                    const double zz = rr / fmax(fabs(apv[m]), SMALL);
                    if (ok[m]) {
                        x[oc[m]] = xv[m] + alpha * pvv[m];
                        r[oc[m]] = rr;
                        zo[oc[m]] = zz;
                        acc[0] += rr * rr;
                        acc[1] += rr * zz;
                    }
                }
            }
            block_sum<3>(acc, sh);
            store_run<3>(a.partB, t, nrun, w.stride, acc);
        }
        t += nrun * w.stride;
    }
}

// Number of finished systems in slot 1 (one block).
__global__ __launch_bounds__(BLOCK) void k64_count(Cg64Args a) {
    __shared__ double sh[BLOCK / 64];
    double v[1] = {0.0};
    for (int s = threadIdx.x; s < a.nsys; s += BLOCK) v[0] += a.slot[1][s].done ? 1.0 : 0.0;
    block_sum<1>(v, sh);
    if (threadIdx.x == 0) *a.ndone = (int)v[0];
}

// First-time buffers of the FP64 solver; the caller publishes w only when
// every one of them exists.
static int init_cg64(Batch& B, Cg64Work* w) {
    {
        const size_t n = B.ncells;
        EAFP_TRY_RC(w->x.alloc(n));
        EAFP_TRY_RC(w->r.alloc(n));
        EAFP_TRY_RC(w->z.alloc(n));
        EAFP_TRY_RC(w->q.alloc(n));
        EAFP_TRY_RC(w->p0.alloc(n));
        EAFP_TRY_RC(w->p1.alloc(n));
        EAFP_TRY_RC(w->slot0.alloc(B.nsys));
        EAFP_TRY_RC(w->slot1.alloc(B.nsys));
        EAFP_TRY_RC(w->cst.alloc(B.nsys));
        EAFP_TRY_RC(w->ndone.alloc(1));
        EAFP_TRY(hipHostMalloc((void**)&w->ndone_host, 2 * sizeof(int), hipHostMallocDefault));
        for (auto& e : w->ev) EAFP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        EAFP_TRY(hipEventCreate(&w->t0));
        EAFP_TRY(hipEventCreate(&w->t1));
        EAFP_TRY(hipMemset(w->p0.p, 0, n * sizeof(double)));
        EAFP_TRY(hipMemset(w->p1.p, 0, n * sizeof(double)));
    }
    return EAFP_OK;
}

static int ensure_cg64(Batch& B) {
    if (!B.cg64) {
        Cg64Work* nw = new Cg64Work();
        const int rc = init_cg64(B, nw);
        if (rc != EAFP_OK) {
            free_cg64(nw);
            return rc;
        }
        B.cg64 = nw;
    }
    Cg64Work* w = B.cg64;
    EAFP_TRY_RC(w->partA.alloc((size_t)B.ntiles));
    EAFP_TRY_RC(w->partB.alloc((size_t)B.ntiles * 3));
    Cg64Args& a = w->args;
    a.tiles = B.tiles.p;
    a.ntiles = B.ntiles;
    a.nsys = B.nsys;
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
    a.x = w->x.p;
    a.r = w->r.p;
    a.z = w->z.p;
    a.q = w->q.p;
    a.p[0] = w->p0.p;
    a.p[1] = w->p1.p;
    a.partA = w->partA.p;
    a.partB = w->partB.p;
    a.slot[0] = w->slot0.p;
    a.slot[1] = w->slot1.p;
    a.state_first = B.nsys > 1 ? 1 : 0;
    a.cst = w->cst.p;
    a.ndone = w->ndone.p;
    return EAFP_OK;
}

template <int CPT>
static void launch_spmv64(Batch& B, int par) {
    k64_spmv<CPT><<<B.grid, BLOCK, 0, B.stream>>>(B.cg64->args, par);
}
template <int CPT>
static void launch_update64(Batch& B, int par) {
    k64_update<CPT><<<B.grid, BLOCK, 0, B.stream>>>(B.cg64->args, par);
}
template <int CPT>
static void launch_init64_t(Batch& B) {
    k64_init<CPT><<<B.grid, BLOCK, 0, B.stream>>>(B.cg64->args);
}

#define EAFP_CPT_DISPATCH(fn, ...)                  \
    switch (B.cpt) {                                \
        case 1: fn<1>(__VA_ARGS__); break;          \
        case 4: fn<4>(__VA_ARGS__); break;          \
        case 8: fn<8>(__VA_ARGS__); break;          \
        default: fn<2>(__VA_ARGS__); break;         \
    }

// The graph of `chunk` iterations for every system, looked up in the
// workspace's cache by chunk, launch shape and every buffer its kernels
// name; captured and instantiated on a miss.
static int cg64_graph(Batch& B, int chunk) {
    Cg64Work* w = B.cg64;
    const int key[GraphCache<Cg64Args>::kKeys] = {chunk, 0, 0, 0};
    w->exec = w->graphs.find(key, B.tiling_gen, w->args);
    if (w->exec) return EAFP_OK;
    EAFP_TRY(hipStreamSynchronize(B.stream));   // no earlier launch of a graph that make_room may drop
    w->graphs.make_room(B.tiling_gen);
    hipGraph_t graph = nullptr;
    EAFP_TRY(hipStreamBeginCapture(B.stream, hipStreamCaptureModeThreadLocal));
    for (int c = 0; c < chunk; ++c) {
        EAFP_CPT_DISPATCH(launch_spmv64, B, c & 1)
        EAFP_CPT_DISPATCH(launch_update64, B, c & 1)
    }
    k64_count<<<1, BLOCK, 0, B.stream>>>(w->args);
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

int solve_fp64(Batch& B, int max_iter, double tol, eafp_stats* stats, int chunk) {
    for (int s = 0; s < B.nsys; ++s)
        if (!B.uploaded[s]) return EAFP_ERR_STATE;
    EAFP_TRY_RC(ensure_cg64(B));
    Cg64Work* w = B.cg64;
    if (chunk <= 0) chunk = 16;
    chunk = (chunk + 1) & ~1;    // even: the pv buffer parity restarts every chunk
    EAFP_TRY_RC(cg64_graph(B, chunk));

    std::vector<CgConst> cst(B.nsys);
    for (auto& K : cst) {
        K = CgConst{};
        K.tol = tol;
        K.max_iter = max_iter;
    }
    EAFP_TRY(hipMemcpyAsync(w->cst.p, cst.data(), sizeof(CgConst) * B.nsys, hipMemcpyHostToDevice, B.stream));
    EAFP_TRY(hipEventRecord(w->t0, B.stream));
    EAFP_CPT_DISPATCH(launch_init64_t, B)
    EAFP_TRY(hipGetLastError());

    // Chunks of `chunk` iterations; poll the finished count of the chunk
    // before the one just launched so the queue never drains.  A system is
    // marked finished by the k64_spmv after its last update, so one chunk
    // past max_iter is always enough.
    const long max_chunks = (long)(max_iter > 0 ? max_iter : 0) / chunk + 3;
    long launched = 0;
    for (;;) {
        if (launched < max_chunks) {
            EAFP_TRY(hipGraphLaunch(w->exec, B.stream));
            counter_add(CTR_LAUNCHES, 1);
            const int slot = (int)(launched & 1);
            EAFP_TRY(hipMemcpyAsync(&w->ndone_host[slot], w->ndone.p, sizeof(int), hipMemcpyDeviceToHost, B.stream));
            EAFP_TRY(hipEventRecord(w->ev[slot], B.stream));
            ++launched;
        }
        if (launched >= 2) {
            const int older = (int)(launched & 1);
            EAFP_TRY(hipEventSynchronize(w->ev[older]));
            if (w->ndone_host[older] >= B.nsys) break;
        }
        if (launched >= max_chunks) {
            EAFP_TRY(hipEventSynchronize(w->ev[(launched - 1) & 1]));
            break;
        }
    }
    for (int s = 0; s < B.nsys; ++s) EAFP_TRY_RC(download_x(B, s, w->x.p));
    EAFP_TRY(hipEventRecord(w->t1, B.stream));
    std::vector<CgSlot> st(B.nsys);
    EAFP_TRY(hipMemcpyAsync(st.data(), w->slot1.p, sizeof(CgSlot) * B.nsys, hipMemcpyDeviceToHost, B.stream));
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

// Average durations of the two iteration kernels over `reps` iterations
// on the uploaded systems (kept running: tol = 0), each launch bracketed
// by HIP events on the library stream.
int time_fp64_kernels(Batch& B, int reps, float* ms_spmv, float* ms_update) {
    for (int s = 0; s < B.nsys; ++s)
        if (!B.uploaded[s]) return EAFP_ERR_STATE;
    if (reps < 1) return EAFP_ERR_ARG;
    EAFP_TRY_RC(ensure_cg64(B));
    Cg64Work* w = B.cg64;
    std::vector<CgConst> cst(B.nsys);
    for (auto& K : cst) {
        K = CgConst{};
        K.tol = 0.0;
        K.max_iter = 1 << 30;
    }
    EAFP_TRY(hipMemcpyAsync(w->cst.p, cst.data(), sizeof(CgConst) * B.nsys, hipMemcpyHostToDevice, B.stream));
    EAFP_CPT_DISPATCH(launch_init64_t, B)
    hipEvent_t e[3];
    for (auto& x : e) EAFP_TRY(hipEventCreate(&x));
    float tot_s = 0.f, tot_u = 0.f;
    for (int r = 0; r < reps; ++r) {
        const int par = r & 1;
        EAFP_TRY(hipEventRecord(e[0], B.stream));
        EAFP_CPT_DISPATCH(launch_spmv64, B, par)
        EAFP_TRY(hipEventRecord(e[1], B.stream));
        EAFP_CPT_DISPATCH(launch_update64, B, par)
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
