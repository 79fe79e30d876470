// tdma.hip -- EAF3D's line-by-line TDMA smoother on the GPU: tdma_3d_mpi
// (src/mod_solver_3d.f90:252-366) with its Thomas solve (:22-46), for one
// serial mesh or a batch of independent ones, in the caller's padded Fortran
// layout (bounds (-1:n+2), r fastest).
//
// The reference sweeps lines in a fixed order and each line reads the
// neighbours of the other directions from phi as it stands: the lines before
// it in that order are already updated, the ones after it are not.  For the
// r pass (loops k outer, j inner) line (j,k) therefore needs lines (j-1,k)
// and (j,k-1) done and (j+1,k), (j,k+1) not started; the theta wrap is read
// through the halo copied before the pass, so it never creates a dependence.
// All lines with the same j + k (one anti-diagonal of the line grid) are
// mutually independent: solving the diagonals one after another, the lines of
// a diagonal in parallel, performs the same arithmetic on the same values as
// the sequential loops.  The theta pass (k outer, i inner) and the z pass
// (j outer, i inner) are scheduled the same way over (i,k) and (i,j).
//
// Kernels:
//   k_tdma_pivots  the pivots m(i) and c'(i) of every line of the three
//                  passes: they depend on the coefficients only, so they are
//                  computed once per call for all lines at once, with the
//                  reference's expressions and clamp (|m| < 1e-30 becomes
//                  sign(1e-30, m)), and reused by every sweep
//   k_tdma_sweeps  one work-group per system runs all sweeps: the theta halo
//                  copy, then each pass one diagonal at a time (one thread
//                  per line, a work-group barrier between diagonals).  Per
//                  line: forward elimination d'(i) = (d(i) - a(i) d'(i-1)) /
//                  m(i), back substitution x(i) = d'(i) - c'(i) x(i+1).
// The critical path of one sweep is (nth+nz-1) r lines + (nr+nz-1) theta
// lines + (nr+nth-1) z lines, each a chain of dependent FP64 divisions; a
// single system is latency-bound by design, and independent systems (the
// equations of one time step) fill the GPU side by side.
//
// Built without floating-point contraction, with the reference's operand
// order, so every value of phi is the double the sequential loops compute.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "batch.hpp"
#include "tdma_line.hpp"

namespace eafp {

namespace {

constexpr int TD_BLOCK = 256;          // threads per system in the sweep kernel

using namespace tdma;

__global__ __launch_bounds__(TD_BLOCK) void k_tdma_pivots(const TdSys* __restrict__ sys) {
    const TdSys s = sys[blockIdx.y];
    int p, u, v;
    if (pivot_line(s, (long)blockIdx.x * TD_BLOCK + threadIdx.x, p, u, v)) line_pivots(s, p, u, v);
}

// The serial periodic copy by the whole work-group, one (i,k) column per thread.
__device__ void theta_copy(const TdSys& s) {
    const int ni = s.nr + 4, nk = s.nz + 4;
    for (int t = threadIdx.x; t < ni * nk; t += TD_BLOCK) theta_copy_column(s, t % ni - 1, t / ni - 1);
}

// One pass over the line grid nu x nv (line (u,v)), diagonal w = (u-1)+(v-1)
// at a time, in the order of the reference's loops (v outer, u inner).
template <int P>
__device__ void run_pass(const TdSys& s, const double* __restrict__ const* c, int nu, int nv) {
    for (int w = 0; w <= nu + nv - 2; ++w) {
        int v_lo, v_hi;
        diagonal_range(w, nu, nv, v_lo, v_hi);
        for (int v = v_lo + (int)threadIdx.x; v <= v_hi; v += TD_BLOCK) solve_line<P>(s, c, w + 2 - v, v);
        __syncthreads();
    }
}

__global__ __launch_bounds__(TD_BLOCK) void k_tdma_sweeps(const TdSys* __restrict__ sys, int n_sweep) {
    const TdSys s = sys[blockIdx.x];
    const double* c[TD_NCOEF];
#pragma unroll
    for (int q = 0; q < TD_NCOEF; ++q) c[q] = s.c[q];
    for (int sweep = 0; sweep < n_sweep; ++sweep) {
        theta_copy(s);
        __syncthreads();
        run_pass<0>(s, c, s.nth, s.nz);     // r lines (j,k)
        theta_copy(s);
        __syncthreads();
        run_pass<1>(s, c, s.nr, s.nz);      // theta lines (i,k)
        theta_copy(s);
        __syncthreads();
        run_pass<2>(s, c, s.nr, s.nth);     // z lines (i,j)
    }
}

struct TdWork {
    hipStream_t stream = nullptr;
    DevBuf<double> work;       // pivots and scratch of every system: 7 n each
    DevBuf<double> stage;      // host arrays: 9 n each (8 coefficients, phi)
    DevBuf<TdSys> table;
};

std::mutex g_td_mutex;
TdWork* g_td = nullptr;

}  // namespace

int tdma_solve_batch(int nsys, const eafp_desc* descs, const eafp_tdma_system* sys, int n_sweep, int on_device) {
    if (nsys < 1 || !descs || !sys || n_sweep < 0) return EAFP_ERR_ARG;
    long total = 0, max_lines = 0;
    for (int q = 0; q < nsys; ++q) {
        EAFP_TRY_RC(validate_desc(&descs[q]));
        const eafp_tdma_system& y = sys[q];
        const double* ptrs[9] = {y.aW, y.aE, y.aS, y.aN, y.aB, y.aT, y.aP, y.Su, y.phi};
        for (const double* p : ptrs)
            if (!p) return EAFP_ERR_ARG;
        const eafp_desc& d = descs[q];
        total += (long)(d.nr + 4) * (d.ntheta + 4) * (d.nz + 4);
        max_lines = std::max(max_lines, (long)d.ntheta * d.nz + (long)d.nr * d.nz + (long)d.nr * d.ntheta);
    }
    if (n_sweep == 0) return EAFP_OK;
    std::lock_guard<std::mutex> lock(g_td_mutex);
    if (!g_td) g_td = new TdWork();
    TdWork& w = *g_td;
    if (!w.stream) EAFP_TRY(hipStreamCreateWithFlags(&w.stream, hipStreamNonBlocking));
    if (w.work.n < (size_t)(7 * total)) EAFP_TRY_RC(w.work.alloc((size_t)(7 * total)));
    if (!on_device && w.stage.n < (size_t)(9 * total)) EAFP_TRY_RC(w.stage.alloc((size_t)(9 * total)));
    if (w.table.n < (size_t)nsys) EAFP_TRY_RC(w.table.alloc((size_t)nsys));
    std::vector<TdSys> tab((size_t)nsys);
    long off = 0;
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
        const double* src[TD_NCOEF] = {y.aW, y.aE, y.aS, y.aN, y.aB, y.aT, y.aP, y.Su};
        if (on_device) {
            for (int c = 0; c < TD_NCOEF; ++c) t.c[c] = src[c];
            t.phi = y.phi;
        } else {
            double* base = w.stage.p + 9 * off;
            for (int c = 0; c < TD_NCOEF; ++c) {
                EAFP_TRY(hipMemcpyAsync(base + c * t.n, src[c], sizeof(double) * t.n, hipMemcpyHostToDevice,
                                        w.stream));
                t.c[c] = base + c * t.n;
            }
            t.phi = base + TD_NCOEF * t.n;
            EAFP_TRY(hipMemcpyAsync(t.phi, y.phi, sizeof(double) * t.n, hipMemcpyHostToDevice, w.stream));
        }
        t.piv = w.work.p + 7 * off;
        t.dp = t.piv + 6 * t.n;
        off += t.n;
    }
    EAFP_TRY(hipMemcpyAsync(w.table.p, tab.data(), sizeof(TdSys) * nsys, hipMemcpyHostToDevice, w.stream));
    const dim3 pgrid((unsigned)((max_lines + TD_BLOCK - 1) / TD_BLOCK), (unsigned)nsys);
    k_tdma_pivots<<<pgrid, TD_BLOCK, 0, w.stream>>>(w.table.p);
    k_tdma_sweeps<<<nsys, TD_BLOCK, 0, w.stream>>>(w.table.p, n_sweep);
    EAFP_TRY(hipGetLastError());
    if (!on_device)
        for (int q = 0; q < nsys; ++q)
            EAFP_TRY(hipMemcpyAsync(sys[q].phi, tab[(size_t)q].phi, sizeof(double) * tab[(size_t)q].n,
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
