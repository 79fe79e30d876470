// step.hip -- one whole p' step of solve_pressure_correction
// (src/mod_pressure_3d.f90:127-397) with every intermediate in HBM: the
// assembly writes the p' system and the liquid's face exports into library
// buffers, the solver packs that system where it lies (no staging copy, no
// host round trip), and the post-solve kernels read p' straight from the
// solver's padded iterate.  Host-array callers pay one upload of the fields
// and one download of the results instead of the three steps' transfers.
#include <algorithm>
#include <memory>
#include <mutex>

#include "batch.hpp"

namespace eafp {
int solve_fp64(Batch& B, int max_iter, double tol, eafp_stats* stats, int chunk);
int solve_mixed(Batch& B, const eafp_mixed_opts& o, eafp_stats* stats);

namespace {

// Double-valued fields in the order of their mirror slots (host-array mode).
constexpr int NFD = 24;   // 23 cell fields + liq_room
constexpr int NIO_D = 7;  // in/out doubles: the six velocities and p

struct StepWork {
    int nr = -1, nth = -1, nz = -1;
    std::unique_ptr<Batch> B;   // single-system batch: the p' solve
    DevBuf<double> sys;         // 8 n: aW, aE, aS, aN, aB, aT, aP, Su
    DevBuf<double> exd;         // 6 n: Fs_r/th/z, ac_r/th/z
    DevBuf<int> exi;            // 3 n: link flags of the exports
    // host-array mode only: device mirrors of the caller's arrays
    DevBuf<double> coord;       // r, theta, z
    DevBuf<double> fd;          // NFD n
    DevBuf<int> fi;             // cell_type (n), masks (3 n), outlet (n2)
    DevBuf<double> iod;         // Fc 3 n, then the in/out arrays not aliased to a field (NIO_D n)
    DevBuf<int> ioi;            // 3 n: lk outputs
};

std::mutex g_step_mutex;
StepWork* g_step = nullptr;

int ensure_step(const eafp_desc* d, bool host_mode) {
    if (!g_step) g_step = new StepWork();
    StepWork& w = *g_step;
    const long n = (long)(d->nr + 4) * (d->ntheta + 4) * (d->nz + 4);
    const long n2 = (long)(d->nr + 4) * (d->ntheta + 4);
    if (!w.B || w.nr != d->nr || w.nth != d->ntheta || w.nz != d->nz) {
        w.nr = w.nth = w.nz = -1;   // a failure below leaves the workspace invalid, not stale
        w.B.reset();
        auto nb = std::make_unique<Batch>();
        EAFP_TRY_RC(nb->init(1, d));
        w.B = std::move(nb);
        EAFP_TRY_RC(w.sys.alloc((size_t)n * 8));
        EAFP_TRY_RC(w.exd.alloc((size_t)n * 6));
        EAFP_TRY_RC(w.exi.alloc((size_t)n * 3));
        w.coord.release();
        w.fd.release();
        w.fi.release();
        w.iod.release();
        w.ioi.release();
        w.nr = d->nr;
        w.nth = d->ntheta;
        w.nz = d->nz;
    }
    if (host_mode && !w.fd.p) {
        EAFP_TRY_RC(w.coord.alloc((size_t)(d->nr + d->ntheta + d->nz + 12)));
        EAFP_TRY_RC(w.fd.alloc((size_t)n * NFD));
        EAFP_TRY_RC(w.fi.alloc((size_t)n * 4 + (size_t)n2));
        EAFP_TRY_RC(w.iod.alloc((size_t)n * (3 + NIO_D)));
        EAFP_TRY_RC(w.ioi.alloc((size_t)n * 3));
    }
    return EAFP_OK;
}

long n_interior(const eafp_desc* d) { return (long)d->nr * d->ntheta * d->nz; }

bool valid_mixed(const eafp_mixed_opts& o) {
    return o.tol > 0.0 && o.max_steps >= 1 && o.interval >= 0 && o.coef_format >= EAFP_COEF_FP32 &&
           o.coef_format <= EAFP_COEF_FP16;
}

}  // namespace

int pp_step(const eafp_desc* d, const eafp_pp_fields* f, const eafp_pp_config* cfg, const eafp_pp_step_opts* o,
            double* pp, const eafp_pp_update* io, double* residual, eafp_stats* stats, int on_device) {
    if (!d || !f || !cfg || !o || !pp || !io) return EAFP_ERR_ARG;
    EAFP_TRY_RC(validate_desc(d));
    const double* fdbl[NFD] = {f->vol,    f->Ar,     f->Ath,    f->Az,    f->alpha_l, f->ur_l,  f->uth_l, f->uz_l,
                               f->aPr_l,  f->aPth_l, f->aPz_l,  f->rho_l, f->T_l,     f->alpha_g, f->ur_g, f->uth_g,
                               f->uz_g,   f->aPr_g,  f->aPth_g, f->aPz_g, f->rho_g,   f->T_old_g, f->p,    f->liq_room};
    for (int q = 0; q < NFD - 1; ++q)
        if (!fdbl[q]) return EAFP_ERR_ARG;
    if (!f->r || !f->theta || !f->z || !f->cell_type || !f->outlet) return EAFP_ERR_ARG;
    double* const fc[3] = {io->Fc_r, io->Fc_th, io->Fc_z};
    int* const lko[3] = {io->lk_r, io->lk_th, io->lk_z};
    double* const inout[NIO_D] = {io->ur_l, io->uth_l, io->uz_l, io->ur_g, io->uth_g, io->uz_g, io->p};
    for (int q = 0; q < 3; ++q)
        if (!fc[q] || !lko[q]) return EAFP_ERR_ARG;
    for (double* q : inout)
        if (!q) return EAFP_ERR_ARG;
    if (!(cfg->dt > 0.0)) return EAFP_ERR_ARG;
    if (o->solver != EAFP_STEP_FP64 && o->solver != EAFP_STEP_MIXED) return EAFP_ERR_ARG;
    if (o->solver == EAFP_STEP_MIXED && !valid_mixed(o->mixed)) return EAFP_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return EAFP_ERR_NO_DEVICE;

    std::lock_guard<std::mutex> lock(g_step_mutex);
    EAFP_TRY_RC(ensure_step(d, !on_device));
    StepWork& w = *g_step;
    Batch& B = *w.B;
    const long n = (long)(d->nr + 4) * (d->ntheta + 4) * (d->nz + 4);
    const long n2 = (long)(d->nr + 4) * (d->ntheta + 4);
    const size_t nd = sizeof(double) * n, ni = sizeof(int) * n;

    // The arrays the three steps read and write, on the device.
    eafp_pp_fields fdev = *f;
    eafp_pp_update iodev = *io;
    if (!on_device) {
        const hipMemcpyKind H2D = hipMemcpyHostToDevice;
        hipStream_t st = B.stream;
        const int nr = d->nr, nth = d->ntheta, nz = d->nz;
        double* c = w.coord.p;
        EAFP_TRY(hipMemcpyAsync(c, f->r, sizeof(double) * (nr + 4), H2D, st));
        EAFP_TRY(hipMemcpyAsync(c + nr + 4, f->theta, sizeof(double) * (nth + 4), H2D, st));
        EAFP_TRY(hipMemcpyAsync(c + nr + nth + 8, f->z, sizeof(double) * (nz + 4), H2D, st));
        fdev.r = c;
        fdev.theta = c + nr + 4;
        fdev.z = c + nr + nth + 8;
        const double** fdst[NFD] = {&fdev.vol,    &fdev.Ar,     &fdev.Ath,    &fdev.Az,    &fdev.alpha_l,
                                    &fdev.ur_l,   &fdev.uth_l,  &fdev.uz_l,   &fdev.aPr_l, &fdev.aPth_l,
                                    &fdev.aPz_l,  &fdev.rho_l,  &fdev.T_l,    &fdev.alpha_g, &fdev.ur_g,
                                    &fdev.uth_g,  &fdev.uz_g,   &fdev.aPr_g,  &fdev.aPth_g, &fdev.aPz_g,
                                    &fdev.rho_g,  &fdev.T_old_g, &fdev.p,     &fdev.liq_room};
        for (int q = 0; q < NFD; ++q) {
            if (!fdbl[q]) continue;   // liq_room is optional
            double* m = w.fd.p + (long)q * n;
            EAFP_TRY(hipMemcpyAsync(m, fdbl[q], nd, H2D, st));
            *fdst[q] = m;
        }
        int* ip = w.fi.p;
        EAFP_TRY(hipMemcpyAsync(ip, f->cell_type, ni, H2D, st));
        fdev.cell_type = ip;
        const int* msk[3] = {f->liq_cont, f->pv_active, f->gas_reach};
        const int** mdst[3] = {&fdev.liq_cont, &fdev.pv_active, &fdev.gas_reach};
        for (int q = 0; q < 3; ++q) {
            if (!msk[q]) continue;
            EAFP_TRY(hipMemcpyAsync(ip + (1 + q) * n, msk[q], ni, H2D, st));
            *mdst[q] = ip + (1 + q) * n;
        }
        EAFP_TRY(hipMemcpyAsync(ip + 4 * n, f->outlet, sizeof(int) * n2, H2D, st));
        fdev.outlet = ip + 4 * n;
        // Fc and lk are written in full by the post-solve step: outputs only.
        double** fcd[3] = {&iodev.Fc_r, &iodev.Fc_th, &iodev.Fc_z};
        int** lkd[3] = {&iodev.lk_r, &iodev.lk_th, &iodev.lk_z};
        for (int q = 0; q < 3; ++q) {
            *fcd[q] = w.iod.p + (long)q * n;
            *lkd[q] = w.ioi.p + (long)q * n;
        }
        // An in/out array that is also a field (liq%ur is both) shares the
        // field's mirror: the assembly reads it before the update writes it.
        const double* const alias[NIO_D] = {f->ur_l, f->uth_l, f->uz_l, f->ur_g, f->uth_g, f->uz_g, f->p};
        const int aslot[NIO_D] = {5, 6, 7, 14, 15, 16, 22};
        double** iodst[NIO_D] = {&iodev.ur_l, &iodev.uth_l, &iodev.uz_l, &iodev.ur_g, &iodev.uth_g, &iodev.uz_g,
                                 &iodev.p};
        for (int q = 0; q < NIO_D; ++q) {
            if (inout[q] == alias[q]) {
                *iodst[q] = w.fd.p + (long)aslot[q] * n;
            } else {
                double* m = w.iod.p + (long)(3 + q) * n;
                EAFP_TRY(hipMemcpyAsync(m, inout[q], nd, H2D, st));
                *iodst[q] = m;
            }
        }
        EAFP_TRY(hipStreamSynchronize(st));
    }

    // sh%pp = 0 and the assembly (src/mod_pressure_3d.f90:174-306)
    double* sys[8];
    for (int q = 0; q < 8; ++q) sys[q] = w.sys.p + (long)q * n;
    eafp_pp_exports ex{w.exd.p, w.exd.p + n, w.exd.p + 2 * n, w.exd.p + 3 * n, w.exd.p + 4 * n, w.exd.p + 5 * n,
                       w.exi.p, w.exi.p + n, w.exi.p + 2 * n};
    double res = 0.0;
    EAFP_TRY_RC(pp_assemble(d, &fdev, cfg, sys, &res, nullptr, &ex, 1));

    // the solve from pp = 0 (:316-318), the system packed where it lies
    const double* csys[8];
    for (int q = 0; q < 8; ++q) csys[q] = sys[q];
    const bool mixed = o->solver == EAFP_STEP_MIXED;
    EAFP_TRY_RC(B.upload(0, csys, nullptr, 1));
    // launch shape by solver, as tuned on the 60x120x84 production mesh: the
    // FP64 CG at one cell per thread, the mixed solver at two; four blocks
    // per CU, at most one per tile
    const int cpt = mixed ? 2 : 1;
    if (B.cpt != cpt) {
        const int tiles = (int)((n_interior(d) + (long)BLOCK * cpt - 1) / ((long)BLOCK * cpt));
        EAFP_TRY_RC(B.set_tiling(cpt, std::min(4 * device_cus(), tiles)));
    }
    eafp_stats st{};
    if (mixed)
        EAFP_TRY_RC(solve_mixed(B, o->mixed, &st));
    else
        EAFP_TRY_RC(solve_fp64(B, o->max_iter, o->tol, &st, 0));
    if (st.residual != st.residual) res = st.residual;   // :322

    // the post-solve steps (:324-397), p' read from the solver's iterate
    const double* ppd = B.phi_pad[0]->p;
    EAFP_TRY_RC(pp_correct(d, &fdev, cfg, o->alpha_p, ppd, &ex, &iodev, 1));
    EAFP_TRY_RC(B.download(0, pp, on_device));

    if (!on_device) {
        const hipMemcpyKind D2H = hipMemcpyDeviceToHost;
        hipStream_t s = B.stream;
        for (int q = 0; q < 3; ++q) {
            EAFP_TRY(hipMemcpyAsync(fc[q], w.iod.p + (long)q * n, nd, D2H, s));
            EAFP_TRY(hipMemcpyAsync(lko[q], w.ioi.p + (long)q * n, ni, D2H, s));
        }
        const double* const iosrc[NIO_D] = {iodev.ur_l, iodev.uth_l, iodev.uz_l, iodev.ur_g,
                                            iodev.uth_g, iodev.uz_g, iodev.p};
        for (int q = 0; q < NIO_D; ++q) EAFP_TRY(hipMemcpyAsync(inout[q], iosrc[q], nd, D2H, s));
        EAFP_TRY(hipStreamSynchronize(s));
    }
    if (residual) *residual = res;
    if (stats) *stats = st;
    return EAFP_OK;
}

void step_release() {
    std::lock_guard<std::mutex> lock(g_step_mutex);
    if (g_step) {
        if (g_step->B && g_step->B->stream) (void)hipStreamSynchronize(g_step->B->stream);
        delete g_step;
        g_step = nullptr;
    }
}

}  // namespace eafp
