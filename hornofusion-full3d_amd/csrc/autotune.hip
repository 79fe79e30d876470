// autotune.hip -- launch-shape autotuner for the iteration kernels.
//
// The two inner-iteration kernels of a solver are timed (HIP events on the
// library stream, averaged over `reps` iterations) for every candidate
// shape: cells per thread in {1, 2, 4, 8} (the tile is 256 * cpt cells)
// and blocks per CU in {4, 8, 16} (capped by the tile count) or one block
// per tile.  For the mixed solver the column-marching stencil is then timed
// at 1 .. 16 r-segments.  The fastest shape is kept in the batch; solvers
// pick it up at their next solve.
#include <cfloat>

#include "batch.hpp"

namespace eafp {

int time_fp64_kernels(Batch& B, int reps, float* ms_spmv, float* ms_update);
int time_mixed_kernels(Batch& B, int fmt, int precond, int reps, float* ms_spmv, float* ms_update);

// solver 0: FP64 CG; 1: mixed with the Jacobi inner preconditioner; 2:
// mixed with the z-line inner preconditioner.
int time_kernels(Batch& B, int solver, int fmt, int reps, float* ms_spmv, float* ms_update) {
    if (solver == 0) return time_fp64_kernels(B, reps, ms_spmv, ms_update);
    if (solver == 1 || solver == 2)
        return time_mixed_kernels(B, fmt, solver == 2 ? EAFP_PRECOND_ZLINE : EAFP_PRECOND_JACOBI, reps, ms_spmv,
                                  ms_update);
    return EAFP_ERR_ARG;
}

int autotune(Batch& B, int solver, int fmt, int reps, int* best_cpt, int* best_blocks, float* best_ms) {
    const int cus = device_cus();
    const int cpts[] = {1, 2, 4, 8};
    const int per_cu[] = {4, 8, 16, 0};
    float best = FLT_MAX;
    int bc = B.cpt, bb = B.grid;
    EAFP_TRY_RC(B.set_columns(0));
    for (int cpt : cpts) {
        for (int k : per_cu) {
            EAFP_TRY_RC(B.set_tiling(cpt, 0));
            // k = 0: one block per tile (the hardware dispatcher walks the table)
            const int blocks = k == 0 ? B.ntiles : std::min(B.ntiles, k * cus);
            EAFP_TRY_RC(B.set_tiling(cpt, blocks));
            float a = 0.f, b = 0.f;
            // one untimed pass settles caches and code, then the timed reps
            EAFP_TRY_RC(time_kernels(B, solver, fmt, 2, &a, &b));
            EAFP_TRY_RC(time_kernels(B, solver, fmt, reps, &a, &b));
            if (a + b < best) {
                best = a + b;
                bc = cpt;
                bb = B.grid;
            }
        }
    }
    EAFP_TRY_RC(B.set_columns(0));
    EAFP_TRY_RC(B.set_tiling(bc, bb));
    if (solver == 1 || solver == 2) {
        // the column-marching stencil, where every system qualifies
        int bs = 0;
        for (int k : {1, 2, 4, 6, 8, 10, 12, 16}) {
            if (B.set_columns(k) != EAFP_OK) continue;
            float a = 0.f, b = 0.f;
            EAFP_TRY_RC(time_kernels(B, solver, fmt, 2, &a, &b));
            EAFP_TRY_RC(time_kernels(B, solver, fmt, reps, &a, &b));
            if (a + b < best) {
                best = a + b;
                bs = k;
            }
        }
        EAFP_TRY_RC(B.set_columns(bs));
    }
    if (best_cpt) *best_cpt = bc;
    if (best_blocks) *best_blocks = bb;
    if (best_ms) *best_ms = best;
    return EAFP_OK;
}

}  // namespace eafp
