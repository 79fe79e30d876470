// capi.hip -- extern "C" entry points of libeafp.so (declared in
// include/eafp.h).  Argument checks happen before any HIP call, so invalid
// descriptors are reported the same way with or without a GPU.
#include <atomic>
#include <memory>
#include <mutex>

#include "batch.hpp"

namespace eafp {
int solve_fp64(Batch& B, int max_iter, double tol, eafp_stats* stats, int chunk);
int solve_mixed(Batch& B, const eafp_mixed_opts& o, eafp_stats* stats);
int time_kernels(Batch& B, int solver, int fmt, int reps, float* ms_spmv, float* ms_update);
int autotune(Batch& B, int solver, int fmt, int reps, int* best_cpt, int* best_blocks, float* best_ms);

namespace {
std::atomic<long long> g_ctr[CTR_COUNT];
}

void counter_add(int which, long long v) {
    if (which >= 0 && which < CTR_COUNT) g_ctr[which].fetch_add(v, std::memory_order_relaxed);
}
}  // namespace eafp

struct eafp_batch {
    eafp::Batch B;
};

namespace {

int device_check() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return EAFP_ERR_NO_DEVICE;
    return EAFP_OK;
}

bool same_shape(const eafp_desc& a, const eafp_desc& b) {
    return a.nr == b.nr && a.ntheta == b.ntheta && a.nz == b.nz;
}

// Cached single-system batch of the drop-in, keyed by mesh shape.
std::mutex g_mu;
std::unique_ptr<eafp_batch> g_cache;
eafp_desc g_cache_desc{};

}  // namespace

extern "C" {

const char* eafp_version(void) { return "eafp 0.1 (gfx950)"; }

int eafp_cg_solve(const eafp_desc* desc, const double* aW, const double* aE, const double* aS,
                  const double* aN, const double* aB, const double* aT, const double* aP,
                  const double* Su, double* phi, int max_iter, double tol, double* residual,
                  int* n_iter) {
    int rc = eafp::validate_desc(desc);
    if (rc != EAFP_OK) return rc;
    if (!aW || !aE || !aS || !aN || !aB || !aT || !aP || !Su || !phi || !residual || !n_iter)
        return EAFP_ERR_ARG;
    if ((rc = device_check()) != EAFP_OK) return rc;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_cache || !same_shape(g_cache_desc, *desc)) {
        g_cache.reset();
        auto nb = std::make_unique<eafp_batch>();
        if ((rc = nb->B.init(1, desc)) != EAFP_OK) return rc;
        g_cache = std::move(nb);
        g_cache_desc = *desc;
    }
    eafp::Batch& B = g_cache->B;
    const double* arr[8] = {aW, aE, aS, aN, aB, aT, aP, Su};
    if ((rc = B.upload(0, arr, phi, 0)) != EAFP_OK) return rc;
    eafp_stats st{};
    if ((rc = eafp::solve_fp64(B, max_iter, tol, &st, 0)) != EAFP_OK) return rc;
    if ((rc = B.download(0, phi, 0)) != EAFP_OK) return rc;
    *residual = st.residual;
    *n_iter = st.n_iter;
    return EAFP_OK;
}

int eafp_release(void) {
    {
        std::lock_guard<std::mutex> lk(g_mu);
        g_cache.reset();
    }
    eafp::step_release();
    eafp::pp_release();
    eafp::tdma_release();
    return EAFP_OK;
}

int eafp_tdma_solve(const eafp_desc* desc, const double* aW, const double* aE, const double* aS, const double* aN,
                    const double* aB, const double* aT, const double* aP, const double* Su, double* phi,
                    int n_sweep, int arrays_on_device) {
    const eafp_tdma_system s = {aW, aE, aS, aN, aB, aT, aP, Su, phi};
    return eafp::tdma_solve_batch(1, desc, &s, n_sweep, arrays_on_device);
}

int eafp_tdma_solve_batch(int nsys, const eafp_desc* descs, const eafp_tdma_system* systems, int n_sweep,
                          int arrays_on_device) {
    return eafp::tdma_solve_batch(nsys, descs, systems, n_sweep, arrays_on_device);
}

int eafp_pp_assemble(const eafp_desc* desc, const eafp_pp_fields* fields, const eafp_pp_config* cfg, double* aW,
                     double* aE, double* aS, double* aN, double* aB, double* aT, double* aP, double* Su,
                     double* residual, double* flux_ref, eafp_pp_exports* exports, int arrays_on_device) {
    double* out[8] = {aW, aE, aS, aN, aB, aT, aP, Su};
    return eafp::pp_assemble(desc, fields, cfg, out, residual, flux_ref, exports, arrays_on_device);
}

int eafp_pp_assemble_noalloc(const eafp_desc* desc, const eafp_pp_fields* fields, const eafp_pp_config* cfg,
                             double* aW, double* aE, double* aS, double* aN, double* aB, double* aT, double* aP,
                             double* Su, double* residual, double* flux_ref, eafp_pp_exports* exports, void* stream) {
    double* out[8] = {aW, aE, aS, aN, aB, aT, aP, Su};
    return eafp::pp_assemble_noalloc(desc, fields, cfg, out, residual, flux_ref, exports,
                                     static_cast<hipStream_t>(stream));
}

int eafp_pp_correct(const eafp_desc* desc, const eafp_pp_fields* fields, const eafp_pp_config* cfg, double alpha_p,
                    const double* pp, const eafp_pp_exports* exports, const eafp_pp_update* io,
                    int arrays_on_device) {
    return eafp::pp_correct(desc, fields, cfg, alpha_p, pp, exports, io, arrays_on_device);
}

int eafp_pp_step(const eafp_desc* desc, const eafp_pp_fields* fields, const eafp_pp_config* cfg,
                 const eafp_pp_step_opts* opts, double* pp, const eafp_pp_update* io, double* residual,
                 eafp_stats* stats, int arrays_on_device) {
    return eafp::pp_step(desc, fields, cfg, opts, pp, io, residual, stats, arrays_on_device);
}

int eafp_batch_create(int nsys, const eafp_desc* descs, eafp_batch** out) {
    if (!out || nsys < 1 || !descs) return EAFP_ERR_ARG;
    *out = nullptr;
    for (int s = 0; s < nsys; ++s) {
        int rc = eafp::validate_desc(&descs[s]);
        if (rc != EAFP_OK) return rc;
    }
    int rc = device_check();
    if (rc != EAFP_OK) return rc;
    auto nb = std::make_unique<eafp_batch>();
    if ((rc = nb->B.init(nsys, descs)) != EAFP_OK) return rc;
    *out = nb.release();
    return EAFP_OK;
}

int eafp_batch_upload(eafp_batch* b, int s, const double* aW, const double* aE, const double* aS,
                      const double* aN, const double* aB, const double* aT, const double* aP,
                      const double* Su, const double* phi0, int arrays_on_device) {
    if (!b) return EAFP_ERR_ARG;
    const double* arr[8] = {aW, aE, aS, aN, aB, aT, aP, Su};
    return b->B.upload(s, arr, phi0, arrays_on_device);
}

int eafp_batch_solve_fp64(eafp_batch* b, int max_iter, double tol, eafp_stats* stats) {
    if (!b) return EAFP_ERR_ARG;
    return eafp::solve_fp64(b->B, max_iter, tol, stats, 0);
}

int eafp_batch_solve_mixed(eafp_batch* b, const eafp_mixed_opts* opts, eafp_stats* stats) {
    if (!b || !opts) return EAFP_ERR_ARG;
    return eafp::solve_mixed(b->B, *opts, stats);
}

int eafp_batch_set_tiling(eafp_batch* b, int cells_per_thread, int blocks) {
    if (!b) return EAFP_ERR_ARG;
    return b->B.set_tiling(cells_per_thread, blocks);
}

int eafp_batch_set_columns(eafp_batch* b, int r_splits) {
    if (!b) return EAFP_ERR_ARG;
    return b->B.set_columns(r_splits);
}

int eafp_batch_get_columns(const eafp_batch* b, int* r_splits) {
    if (!b || !r_splits) return EAFP_ERR_ARG;
    *r_splits = b->B.col_split;
    return EAFP_OK;
}

int eafp_batch_time_kernels(eafp_batch* b, int solver, int coef_format, int reps, float* ms_spmv,
                            float* ms_update) {
    if (!b || reps < 1) return EAFP_ERR_ARG;
    return eafp::time_kernels(b->B, solver, coef_format, reps, ms_spmv, ms_update);
}

int eafp_batch_autotune(eafp_batch* b, int solver, int coef_format, int reps, int* cells_per_thread,
                        int* blocks, float* ms_iter) {
    if (!b || reps < 1) return EAFP_ERR_ARG;
    return eafp::autotune(b->B, solver, coef_format, reps, cells_per_thread, blocks, ms_iter);
}

int eafp_batch_download(eafp_batch* b, int s, double* phi, int to_device) {
    if (!b) return EAFP_ERR_ARG;
    return b->B.download(s, phi, to_device);
}

int eafp_batch_info(const eafp_batch* b, int* grid, int* tile_cells, long* ncells) {
    if (!b) return EAFP_ERR_ARG;
    if (grid) *grid = b->B.grid;
    if (tile_cells) *tile_cells = eafp::BLOCK * b->B.cpt;
    if (ncells) *ncells = b->B.ncells;
    return EAFP_OK;
}

int eafp_batch_components(eafp_batch* b, int s, int* nsingular, int cells[4]) {
    if (!b || !nsingular) return EAFP_ERR_ARG;
    eafp::Batch& B = b->B;
    if (s < 0 || s >= B.nsys || !B.uploaded[s]) return EAFP_ERR_STATE;
    int rc = B.label_components();
    if (rc != EAFP_OK) return rc;
    const eafp::CompInfo& c = B.comp_info[s];
    *nsingular = c.nsingular;
    if (cells)
        for (int q = 0; q < eafp::MAX_COMP; ++q) cells[q] = q < c.ncomp ? c.cells[q] : 0;
    return EAFP_OK;
}

int eafp_batch_destroy(eafp_batch* b) {
    delete b;
    return EAFP_OK;
}

int eafp_get_counters(eafp_counters* out) {
    if (!out) return EAFP_ERR_ARG;
    out->uploads = eafp::g_ctr[eafp::CTR_UPLOADS].load();
    out->labelled = eafp::g_ctr[eafp::CTR_LABELLED].load();
    out->graph_captures = eafp::g_ctr[eafp::CTR_CAPTURES].load();
    out->graph_launches = eafp::g_ctr[eafp::CTR_LAUNCHES].load();
    return EAFP_OK;
}

}  // extern "C"
