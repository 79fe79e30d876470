// tdma.hip -- EAF3D's line-by-line TDMA smoother on the GPU: tdma_3d_mpi
// (src/mod_solver_3d.f90:252-366) with its Thomas solve (:22-46), for one
// serial mesh or a batch of independent ones; the caller's arrays are in the
// padded Fortran layout (bounds (-1:n+2), r fastest).
//
// The reference sweeps lines in a fixed order and each line reads the
// neighbours of the other directions from phi as it stands: the lines before
// it in that order are already updated, the ones after it are not.  For the
// r pass (loops k outer, j inner) line (j,k) therefore needs lines (j-1,k)
// and (j,k-1) done and (j+1,k), (j,k+1) not started; the theta wrap is read
// through the halo copied before the pass.  All lines with the same j + k
// (one anti-diagonal of the line grid) are mutually independent: solving the
// diagonals one after another, the lines of a diagonal in parallel, performs
// the same arithmetic on the same values as the sequential loops.  The theta
// pass (k outer, i inner) and the z pass (j outer, i inner) are scheduled the
// same way over (i,k) and (i,j).
//
// Each pass keeps its own copy of phi and of its coefficients in a
// diagonal-major layout (tdma_line.hpp): for every element index the lines of
// one diagonal are contiguous, so a work-group loads a whole diagonal
// coalesced whatever the pass direction.  Kernels:
//   k_tdma_pack    once per call, one work-group per tile of 16^3 cells:
//                  the right-hand-side coefficients, a(e), aP and a_plus of
//                  every pass, and pass 0's phi, from the caller's arrays
//                  (read in their own order) into the passes' layouts
//                  (written in theirs), through LDS
//   k_tdma_setup   once per call, one thread per line of each pass: the
//                  pivots m(e) and c'(e) of the forward elimination (they
//                  depend on the coefficients only), the chain kernel's
//                  records, the end terms and the wall halo lines
//   k_tdma_prep<P> and k_tdma_chain<P>: the pass on one work-group per system,
//                  the lines of a diagonal on the lanes of the chain waves
//                  (the schedule is described in tdma_line.hpp); it runs a
//                  pass wherever chain_plan admits it and the pass's pivots
//                  are in its division's band
//   k_tdma_pass<P> the other passes: one 1024-thread work-group per system: per
//                  diagonal, all threads stage d(e) (the right-hand side from
//                  the neighbour lines, whose indices follow from the
//                  diagonal starts), a(e), m(e) and c'(e) of the diagonal's
//                  lines into 160 KiB of LDS with many loads in flight, then
//                  one thread per line runs the dependent Thomas recurrences
//                  from LDS, a chunk of elements read ahead, leaving x in
//                  LDS; then all threads copy the diagonal's x to phi,
//                  coalesced; barriers between the three steps.  A diagonal too long for the LDS is taken in
//                  groups of lines.
//   k_tdma_move    moves phi between the passes' layouts (and into the
//                  caller's array at the end), a tile of 16^3 cells per
//                  work-group through LDS, as k_tdma_pack
//   k_tdma_wall_halo  the theta halo planes of the caller's array on the r/z
//                  wall columns
// A single system is latency-bound by design: one CU walks the
// (nth+nz-1) + (nr+nz-1) + (nr+nth-1) diagonals of a sweep, each a chain of
// dependent FP64 steps; independent systems (the equations of one time
// step) run on other CUs side by side.
//
// Built without floating-point contraction, with the reference's operand
// order, so every value of phi is the double the sequential loops compute.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "batch.hpp"
#include "tdma_line.hpp"

namespace eafp {

namespace {

constexpr int TD_BLOCK = 256;          // threads per work-group of the all-CU kernels
constexpr int TD_PASS_BLOCK = 1024;    // threads of a pass work-group (one per system)

using namespace tdma;

// One tile of one system per work-group (pack_tile_*): each source array is
// staged in one of two LDS buffers while the other one's values are stored,
// one barrier per source.
constexpr int TD_PACK_BLOCK = 512;
__global__ __launch_bounds__(TD_PACK_BLOCK) void k_tdma_pack(const TdSys* __restrict__ sys) {
    __shared__ double buf[2][TD_PT3];
    __shared__ int order[3][TD_PT * TD_PT], lidx[3][TD_PT * TD_PT];
    const TdSys& s = sys[blockIdx.y];
    PackTile t;
    if (!pack_tile_of(s, blockIdx.x, t)) return;     // the whole work-group
    const int tid = (int)threadIdx.x;
    for (int P = 0; P < 3; ++P) pack_tile_order(s, P, t, order[P], lidx[P], tid, TD_PACK_BLOCK);
    pack_tile_load(s, t, 0, buf[0], tid, TD_PACK_BLOCK);
    __syncthreads();
    for (int src = 0; src < TD_PSRC; ++src) {
        if (src + 1 < TD_PSRC) pack_tile_load(s, t, src + 1, buf[(src + 1) & 1], tid, TD_PACK_BLOCK);
        for (int P = 0; P < 3; ++P) pack_tile_store(s, P, t, src, buf[src & 1], order[P], lidx[P], tid, TD_PACK_BLOCK);
        __syncthreads();
    }
}

__global__ __launch_bounds__(TD_BLOCK) void k_tdma_setup(const TdSys* __restrict__ sys) {
    const TdSys& s = sys[blockIdx.y];
    long t = (long)blockIdx.x * TD_BLOCK + threadIdx.x;
    for (int P = 0; P < 3; ++P) {
        const TdPass& p = s.ps[P];
        if (t < p.NL) {
            int u, v;
            line_at(p, (int)t, u, v);
            setup_line(s, P, u, v);
            return;
        }
        t -= p.NL;
    }
}

template <int P>
__global__ __launch_bounds__(TD_PASS_BLOCK) void k_tdma_pass(const TdSys* __restrict__ sys) {
    __shared__ double sd[TD_LDS_ELEMS], sa[TD_LDS_ELEMS], sm[TD_LDS_ELEMS], sc[TD_LDS_ELEMS];
    const TdSys& s = sys[blockIdx.x];
    const TdPass& p = s.ps[P];
    if (p.rec && !s.off_band[P]) return;     // k_tdma_chain runs this pass
    if (P != 1) {     // the halo copy before the pass: snapshot of the lines the wrap reads
        const long nsnap = (long)p.n * (P == 0 ? p.nv : p.nu);
        for (long t = threadIdx.x; t < nsnap; t += TD_PASS_BLOCK) snapshot_value(p, P, t);
        __syncthreads();
    }
    const int gmax = group_lines(p.n);
    for (int w = 0; w <= p.nu + p.nv - 2; ++w) {
        const DiagCtx dc = diag_ctx(p, w);
        int v_lo, v_hi;
        diagonal_range(w, p.nu, p.nv, v_lo, v_hi);
        const int m = v_hi - v_lo + 1, l0 = p.dstart[w];
        for (int g0 = 0; g0 < m; g0 += gmax) {
            const int gm = min(gmax, m - g0), cnt = gm * p.n;
#pragma unroll 2
            for (int idx = threadIdx.x; idx < cnt; idx += TD_PASS_BLOCK) {
                const int q = idx % gm, e = idx / gm;
                const int v = v_lo + g0 + q, u = w + 2 - v;
                int nb[4];
                neighbours_in(p, P, dc, u, v, nb);
                stage_elem<P>(p, l0 + g0 + q, e, nb, sd, sa, sm, sc, idx);
            }
            __syncthreads();
            for (int q = threadIdx.x; q < gm; q += TD_PASS_BLOCK) solve_staged(p, l0 + g0 + q, sd, sa, sm, sc, q, gm);
            __syncthreads();
            for (int idx = threadIdx.x; idx < cnt; idx += TD_PASS_BLOCK) store_staged(p, l0 + g0, gm, sd, idx);
            __syncthreads();
        }
    }
}

// The per-pass products of the chain kernel's records (prep_elem), one
// thread per element of every line.
template <int P>
__global__ __launch_bounds__(TD_BLOCK) void k_tdma_prep(const TdSys* __restrict__ sys) {
    const TdPass& p = sys[blockIdx.y].ps[P];
    const long t = (long)blockIdx.x * TD_BLOCK + threadIdx.x;
    if (t < (long)p.n * p.NL) prep_elem<P>(p, (int)(t % p.NL), (int)(t / p.NL));
}

// One forward chunk f for a helper lane: issue the loads of chunk f+2 into
// `next` (at the end instead with three entries per lane), wait at the
// barriers of the phases up to the one before f's, read the same chunk of
// the diagonal before from the x slots (the chain overwrites it next), fill
// f's ring slot from `use` (loaded two chunks ago), store the copy to phi,
// and close the phase.  Past the pass's last chunk
// (`live` false, the same for every lane) and on the first diagonal the
// copies go to the sink, and past the last chunk nothing is filled and no
// barrier passed.  Every call issues the same stores and loads, so the
// compiler's waits for `use` count exactly the accesses issued after it; a
// wave whose entries all lie past the chunk (h0, its first lane's h, is
// uniform) skips the fill and the copy's LDS read.
template <int P, int IPL>
__device__ __forceinline__ void helper_chunk(const TdPass& p, const TdLds& L, long& X, TdTask& f, TdTask& fl, int h,
                                             int h0, int HL, bool live, const TdPre (&use)[IPL], TdPre (&next)[IPL]) {
    const long ph = chain_prep_phase(p, f);
    // chunk fl's loads first: they depend on nothing this phase does, and
    // they drain through the memory pipeline while the fill and the copy run
    // (with three entries per lane the registers do not allow it)
    if constexpr (IPL < 3) {
#pragma unroll
        for (int i = 0; i < IPL; ++i) chain_load<P>(p, fl, h + i * HL, next[i]);
    }
    if (live)
        for (; X < ph; ++X) __syncthreads();
    // the copy's LDS reads before the fill, its stores after it
    const bool real = live && f.w > 0;
    const int pcnt = real ? f.ce * f.pm : 0;
    double cx[IPL];
    gdptr cd[IPL];
#pragma unroll
    for (int i = 0; i < IPL; ++i) {
        cx[i] = 0.0;
        cd[i] = p.sink + ((h + i * HL) & (TDC_SINK - 1));
        if (h0 + i * HL < pcnt)
            chain_copy_read(p, L, true, f.w - 1, f.pv_lo, f.pm, f.pmM, f.pl0, f.e0, f.ce, h + i * HL, cx[i], cd[i]);
    }
    if (live) {
        const int cnt = f.ce * f.m;
#pragma unroll
        for (int i = 0; i < IPL; ++i)
            if (h0 + i * HL < cnt) chain_prepare<P>(p, L, f, (int)((ph + 1) & 1), h + i * HL, use[i]);
    }
#pragma unroll
    for (int i = 0; i < IPL; ++i) *cd[i] = cx[i];
    if constexpr (IPL >= 3) {
#pragma unroll
        for (int i = 0; i < IPL; ++i) chain_load<P>(p, fl, h + i * HL, next[i]);
    }
    if (live) {
        __syncthreads();
        ++X;
    }
    chain_advance_fwd(p, f);
    chain_advance_fwd(p, fl);
}

// One pass of one system per work-group (the schedule is described in
// tdma_line.hpp): the first NW waves run the recurrences at raised wave
// priority, the other waves feed them, IPL ring entries per helper lane and
// chunk (at least the pass's p.IPL).  Every wave passes the same number of
// barriers, one per phase and one after the x slots' initial ones; the
// branch between the two roles is uniform per wave.
template <int P, int IPL>
__global__ __launch_bounds__(TDC_BLOCK) void k_tdma_chain(const TdSys* __restrict__ sys) {
    __shared__ __align__(16) double lds[TDC_LDS];
    if (sys[blockIdx.x].off_band[P]) return;    // a pivot outside the band: k_tdma_pass runs this pass
    const TdPass p = sys[blockIdx.x].ps[P];     // by value: its fields stay in scalar registers
    const TdLds L = chain_lds(p, lds);
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const long nph = chain_phases(p);
    chain_init_slots(p, L, (int)threadIdx.x, TDC_BLOCK);
    __syncthreads();
    if (wave < p.NW) {
        __builtin_amdgcn_s_setprio(3);
        const int j = (int)threadIdx.x;
        double dp = 0.0;
        TdTask t = chain_first(p);
        for (long X = 0; X < nph; ++X) {
            if (t.kind == 1) chain_forward(p, L, t, (int)(X & 1), j, dp);
            else if (t.kind == 2) chain_backward(p, L, t, j);
            __syncthreads();
            chain_advance(p, t);
        }
        __builtin_amdgcn_s_setprio(0);
    } else {
        const int HL = TDC_BLOCK - 64 * p.NW, h = (int)threadIdx.x - 64 * p.NW, h0 = 64 * (wave - p.NW);
        const long nf = (long)(p.nu + p.nv - 1) * p.K;     // forward chunks of the pass
        TdPre r0[IPL], r1[IPL], r2[IPL];
        TdTask f = chain_first_fwd(p), fl = f;
#pragma unroll
        for (int i = 0; i < IPL; ++i) chain_load<P>(p, fl, h + i * HL, r0[i]);
        chain_advance_fwd(p, fl);
#pragma unroll
        for (int i = 0; i < IPL; ++i) chain_load<P>(p, fl, h + i * HL, r1[i]);
        chain_advance_fwd(p, fl);
        long X = 0;
        for (long q = 0; q < nf; q += 3) {
            helper_chunk<P, IPL>(p, L, X, f, fl, h, h0, HL, true, r0, r2);
            helper_chunk<P, IPL>(p, L, X, f, fl, h, h0, HL, q + 1 < nf, r1, r0);
            helper_chunk<P, IPL>(p, L, X, f, fl, h, h0, HL, q + 2 < nf, r2, r1);
        }
        for (; X < nph; ++X) __syncthreads();
    }
    // the last diagonal's x, after the last barrier
    TdTask t = chain_first(p);
    for (int w = 0; w < p.nu + p.nv - 2; ++w) chain_next_diagonal(p, t);
    const int cnt = p.n * t.m;
    for (int it = (int)threadIdx.x; it < cnt; it += TDC_BLOCK)
        chain_copy_out(p, L, true, t.w, t.v_lo, t.m, t.mM, t.l0, 0, p.n, it);
}

template <int P>
void launch_chain(int ipl, int nsys, hipStream_t st, const TdSys* tp) {
    if (ipl <= 1) k_tdma_chain<P, 1><<<nsys, TDC_BLOCK, 0, st>>>(tp);
    else if (ipl == 2) k_tdma_chain<P, 2><<<nsys, TDC_BLOCK, 0, st>>>(tp);
    else k_tdma_chain<P, 3><<<nsys, TDC_BLOCK, 0, st>>>(tp);
}

// One tile of one system per work-group (move_tile_*).
template <int P, int Q>
__global__ __launch_bounds__(TD_PACK_BLOCK) void k_tdma_move(const TdSys* __restrict__ sys) {
    __shared__ double buf[TD_PT3];
    __shared__ int order[2][TD_PT * TD_PT], lidx[2][TD_PT * TD_PT];
    const TdSys& s = sys[blockIdx.y];
    PackTile t;
    if (!pack_tile_of(s, blockIdx.x, t)) return;     // the whole work-group
    const int tid = (int)threadIdx.x;
    pack_tile_order(s, P, t, order[0], lidx[0], tid, TD_PACK_BLOCK);
    if (Q < 3) pack_tile_order(s, Q, t, order[1], lidx[1], tid, TD_PACK_BLOCK);
    __syncthreads();
    move_tile_gather(s, P, t, buf, order[0], lidx[0], tid, TD_PACK_BLOCK);
    __syncthreads();
    move_tile_scatter(s, P, Q, t, buf, order[1], lidx[1], tid, TD_PACK_BLOCK);
}

__global__ __launch_bounds__(TD_BLOCK) void k_tdma_wall_halo(const TdSys* __restrict__ sys) {
    const TdSys& s = sys[blockIdx.y];
    const long t = (long)blockIdx.x * TD_BLOCK + threadIdx.x;
    if (t < (long)(s.nr + 4) * (s.nz + 4)) wall_halo_column(s, t);
}

struct TdWork {
    hipStream_t stream = nullptr;
    DevBuf<double> work;       // packed arrays and pass copies of phi of every system
    DevBuf<int> iwork;         // diagonal starts
    DevBuf<double> stage;      // host arrays: 9 padded arrays per system
    DevBuf<TdSys> table;
};

std::mutex g_td_mutex;
TdWork* g_td = nullptr;

// Line grid of pass P: (nu, nv, n).
void pass_shape(const eafp_desc& d, int P, int& nu, int& nv, int& n) {
    if (P == 0) { nu = d.ntheta; nv = d.nz; n = d.nr; }
    else if (P == 1) { nu = d.nr; nv = d.nz; n = d.ntheta; }
    else { nu = d.nr; nv = d.ntheta; n = d.nz; }
}

int halo_lines(int P, int nu, int nv) { return P == 0 ? 2 * nu + nv : P == 1 ? 2 * nv + 2 * nu : 2 * nv + nu; }

}  // namespace

int tdma_solve_batch(int nsys, const eafp_desc* descs, const eafp_tdma_system* sys, int n_sweep, int on_device) {
    if (nsys < 1 || !descs || !sys || n_sweep < 0) return EAFP_ERR_ARG;
    long dwork = 0, iwork = 0, stage = 0, max_lines = 0, max_cells = 0, max_cols = 0, max_tiles = 0;
    // A pass runs the chain kernel when every system admits it (chain_plan);
    // EAFP_TDMA_KERNEL=pass keeps k_tdma_pass for every pass.
    const char* force = std::getenv("EAFP_TDMA_KERNEL");
    bool chain[3] = {true, true, true};
    int ipl[3] = {1, 1, 1};     // the chain kernel's ring entries per helper lane: the most any system needs
    if (force && std::strcmp(force, "pass") == 0) chain[0] = chain[1] = chain[2] = false;
    for (int q = 0; q < nsys; ++q)
        for (int P = 0; P < 3; ++P) {
            TdPass t{};
            pass_shape(descs[q], P, t.nu, t.nv, t.n);
            t.NL = t.nu * t.nv;
            t.NH = halo_lines(P, t.nu, t.nv);
            t.NLX = t.NL + t.NH;
            if (!chain_plan(t, 1 << 20)) chain[P] = false;
            else ipl[P] = std::max(ipl[P], t.IPL);
        }
    for (int q = 0; q < nsys; ++q) {
        EAFP_TRY_RC(validate_desc(&descs[q]));
        const eafp_tdma_system& y = sys[q];
        const double* ptrs[9] = {y.aW, y.aE, y.aS, y.aN, y.aB, y.aT, y.aP, y.Su, y.phi};
        for (const double* p : ptrs)
            if (!p) return EAFP_ERR_ARG;
        const eafp_desc& d = descs[q];
        const long ncells = (long)d.nr * d.ntheta * d.nz;
        if (std::max(d.nr, std::max(d.ntheta, d.nz)) > TD_LDS_ELEMS) return EAFP_ERR_UNSUPPORTED;
        long lines = 0;
        for (int P = 0; P < 3; ++P) {
            int nu, nv, n;
            pass_shape(d, P, nu, nv, n);
            const long NL = (long)nu * nv, NLX = NL + halo_lines(P, nu, nv);
            dwork += PK_N * ncells + LN_N * NL + (long)n * NLX + (chain[P] ? (RC_REC + 2) * ncells + TDC_SINK + 2 : 0);
            iwork += nu + nv - 1 + (chain[P] ? NL : 0);
            lines += NL;
        }
        iwork += 3;     // off_band flags
        stage += 9 * (long)(d.nr + 4) * (d.ntheta + 4) * (d.nz + 4);
        max_lines = std::max(max_lines, lines);
        max_cells = std::max(max_cells, ncells);
        max_cols = std::max(max_cols, (long)(d.nr + 4) * (d.nz + 4));
        max_tiles = std::max(max_tiles, (long)((d.nr + TD_PT - 1) / TD_PT) * ((d.ntheta + TD_PT - 1) / TD_PT) *
                                            ((d.nz + TD_PT - 1) / TD_PT));
    }
    if (n_sweep == 0) return EAFP_OK;
    std::lock_guard<std::mutex> lock(g_td_mutex);
    if (!g_td) g_td = new TdWork();
    TdWork& w = *g_td;
    if (!w.stream) EAFP_TRY(hipStreamCreateWithFlags(&w.stream, hipStreamNonBlocking));
    if (w.work.n < (size_t)dwork) EAFP_TRY_RC(w.work.alloc((size_t)dwork));
    if (w.iwork.n < (size_t)iwork) EAFP_TRY_RC(w.iwork.alloc((size_t)iwork));
    if (!on_device && w.stage.n < (size_t)stage) EAFP_TRY_RC(w.stage.alloc((size_t)stage));
    if (w.table.n < (size_t)nsys) EAFP_TRY_RC(w.table.alloc((size_t)nsys));
    std::vector<TdSys> tab((size_t)nsys);
    std::vector<int> ihost((size_t)iwork);
    double* dcur = w.work.p;
    int* icur = w.iwork.p;
    long ioff = 0, soff = 0;
    for (int q = 0; q < nsys; ++q) {
        const eafp_desc& d = descs[q];
        const eafp_tdma_system& y = sys[q];
        TdSys& t = tab[(size_t)q];
        t.nr = d.nr;
        t.nth = d.ntheta;
        t.nz = d.nz;
        t.pad_ = 0;
        t.sy = d.nr + 4;
        t.sz = (long)(d.nr + 4) * (d.ntheta + 4);
        t.n = t.sz * (d.nz + 4);
        t.ncells = (long)d.nr * d.ntheta * d.nz;
        const double* src[TD_NCOEF] = {y.aW, y.aE, y.aS, y.aN, y.aB, y.aT, y.aP, y.Su};
        if (on_device) {
            for (int c = 0; c < TD_NCOEF; ++c) t.c[c] = (cgdptr)src[c];
            t.home = (gdptr)y.phi;
        } else {
            double* base = w.stage.p + soff;
            for (int c = 0; c < TD_NCOEF; ++c) {
                EAFP_TRY(hipMemcpyAsync(base + c * t.n, src[c], sizeof(double) * t.n, hipMemcpyHostToDevice,
                                        w.stream));
                t.c[c] = (cgdptr)(base + c * t.n);
            }
            t.home = (gdptr)(base + TD_NCOEF * t.n);
            EAFP_TRY(hipMemcpyAsync((void*)t.home, y.phi, sizeof(double) * t.n, hipMemcpyHostToDevice, w.stream));
            soff += 9 * t.n;
        }
        t.off_band = (decltype(t.off_band))(icur + ioff);   // three zeros of ihost: cleared by the upload below
        ioff += 3;
        for (int P = 0; P < 3; ++P) {
            TdPass& p = t.ps[P];
            pass_shape(d, P, p.nu, p.nv, p.n);
            p.NL = p.nu * p.nv;
            p.NH = halo_lines(P, p.nu, p.nv);
            p.NLX = p.NL + p.NH;
            p.pk = (gdptr)dcur;
            dcur += PK_N * t.ncells;
            p.ln = (gdptr)dcur;
            dcur += LN_N * (long)p.NL;
            p.phi = (gdptr)dcur;
            dcur += (long)p.n * p.NLX;
            p.rec = p.sink = p.pp = nullptr;
            p.uv = nullptr;
            p.C = p.K = p.ML = p.by_u = p.NW = p.IPL = 0;
            if (chain[P]) {
                if (!chain_plan(p, 1 << 20)) return EAFP_ERR_UNSUPPORTED;   // admitted above: cannot happen
                dcur += (dcur - w.work.p) & 1;      // 16-byte records and pairs
                p.rec = (gdptr)dcur;
                dcur += RC_REC * t.ncells;
                p.pp = (gdptr)dcur;
                dcur += 2 * t.ncells;
                p.sink = (gdptr)dcur;
                dcur += TDC_SINK;
                p.uv = (decltype(p.uv))(icur + ioff);
                ioff += p.NL;
            }
            p.dstart = (cgiptr)(icur + ioff);
            int acc = 0;
            for (int wd = 0; wd <= p.nu + p.nv - 2; ++wd) {
                int lo, hi;
                diagonal_range(wd, p.nu, p.nv, lo, hi);
                ihost[(size_t)ioff + wd] = acc;
                acc += hi - lo + 1;
            }
            ioff += p.nu + p.nv - 1;
        }
    }
    EAFP_TRY(hipMemcpyAsync(w.iwork.p, ihost.data(), sizeof(int) * iwork, hipMemcpyHostToDevice, w.stream));
    EAFP_TRY(hipMemcpyAsync(w.table.p, tab.data(), sizeof(TdSys) * nsys, hipMemcpyHostToDevice, w.stream));
    const unsigned ny = (unsigned)nsys;
    const dim3 g_lines((unsigned)((max_lines + TD_BLOCK - 1) / TD_BLOCK), ny);
    const dim3 g_cells((unsigned)((max_cells + TD_BLOCK - 1) / TD_BLOCK), ny);
    const dim3 g_cols((unsigned)((max_cols + TD_BLOCK - 1) / TD_BLOCK), ny);
    const TdSys* tp = w.table.p;
    const dim3 g_tiles((unsigned)max_tiles, ny);
    k_tdma_pack<<<g_tiles, TD_PACK_BLOCK, 0, w.stream>>>(tp);
    k_tdma_setup<<<g_lines, TD_BLOCK, 0, w.stream>>>(tp);
    for (int sweep = 0; sweep < n_sweep; ++sweep) {
        // with the chain kernel planned, k_tdma_pass still runs the systems whose
        // pivots leave its band (each kernel returns at once for the others)
        if (chain[0]) {
            k_tdma_prep<0><<<g_cells, TD_BLOCK, 0, w.stream>>>(tp);
            launch_chain<0>(ipl[0], nsys, w.stream, tp);
        }
        k_tdma_pass<0><<<nsys, TD_PASS_BLOCK, 0, w.stream>>>(tp);
        k_tdma_move<0, 1><<<g_tiles, TD_PACK_BLOCK, 0, w.stream>>>(tp);
        if (chain[1]) {
            k_tdma_prep<1><<<g_cells, TD_BLOCK, 0, w.stream>>>(tp);
            launch_chain<1>(ipl[1], nsys, w.stream, tp);
        }
        k_tdma_pass<1><<<nsys, TD_PASS_BLOCK, 0, w.stream>>>(tp);
        k_tdma_move<1, 2><<<g_tiles, TD_PACK_BLOCK, 0, w.stream>>>(tp);
        if (chain[2]) {
            k_tdma_prep<2><<<g_cells, TD_BLOCK, 0, w.stream>>>(tp);
            launch_chain<2>(ipl[2], nsys, w.stream, tp);
        }
        k_tdma_pass<2><<<nsys, TD_PASS_BLOCK, 0, w.stream>>>(tp);
        if (sweep + 1 < n_sweep) k_tdma_move<2, 0><<<g_tiles, TD_PACK_BLOCK, 0, w.stream>>>(tp);
        else k_tdma_move<2, 3><<<g_tiles, TD_PACK_BLOCK, 0, w.stream>>>(tp);
    }
    k_tdma_wall_halo<<<g_cols, TD_BLOCK, 0, w.stream>>>(tp);
    EAFP_TRY(hipGetLastError());
    if (!on_device)
        for (int q = 0; q < nsys; ++q)
            EAFP_TRY(hipMemcpyAsync(sys[q].phi, (const void*)tab[(size_t)q].home, sizeof(double) * tab[(size_t)q].n,
                                    hipMemcpyDeviceToHost, w.stream));
    EAFP_TRY(hipStreamSynchronize(w.stream));
    return EAFP_OK;
}

void tdma_release() {
    std::lock_guard<std::mutex> lock(g_td_mutex);
    if (g_td) {
        if (g_td->stream) (void)hipStreamSynchronize(g_td->stream);
        if (g_td->stream) (void)hipStreamDestroy(g_td->stream);
        delete g_td;
        g_td = nullptr;
    }
}

}  // namespace eafp
