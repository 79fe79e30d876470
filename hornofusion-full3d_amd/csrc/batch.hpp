// batch.hpp -- host-side container of one or more p' systems resident in HBM.
//
// A Batch owns, for every system, the FP64 operand set of cg_3d_mpi
// (aW, aE, aS, aN, aB, aT, aP and Su, src/mod_solver_3d.f90:485-496) in the
// compact z-fastest layout, the initial phi, and a padded FP64 copy of phi
// used to hand the result back in the caller's (-1:n+2) Fortran layout
// without touching the caller's r/z halos.  Solvers (cg64.hip, mixed.hip)
// attach their work vectors to the Batch.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "eafp_common.hpp"
#include "../../include/eafp.h"

namespace eafp {

#define EAFP_TRY(expr)                                  \
    do {                                                \
        hipError_t e_ = (expr);                         \
        if (e_ != hipSuccess) return (int)e_;           \
    } while (0)

#define EAFP_TRY_RC(expr)                               \
    do {                                                \
        int rc_ = (expr);                               \
        if (rc_ != EAFP_OK) return rc_;                 \
    } while (0)

// Device buffer with RAII release.
template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    int alloc(size_t count) {
        if (count == n && p) return EAFP_OK;
        release();
        if (count == 0) return EAFP_OK;
        hipError_t e = hipMalloc((void**)&p, count * sizeof(T));
        if (e != hipSuccess) { p = nullptr; return (int)e; }
        n = count;
        return EAFP_OK;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

// Singular (pure-Neumann) connected components of one system, found on the
// host at upload: the mixed solver projects its residuals and its result to
// zero mean on each of them.
struct CompInfo {
    int ncomp = 0;                 // projected components (<= MAX_COMP)
    int overflow = 0;              // more singular components than MAX_COMP
    double inv_cnt[MAX_COMP] = {}; // 1 / cells of each
};

int analyze_components(const SysGeom& g, const double* const padded[7], unsigned char* slot, CompInfo& info);

struct Cg64Work;     // cg64.hip
struct MixedWork;    // mixed.hip

enum { C_AW = 0, C_AE, C_AS, C_AN, C_AB, C_AT, C_AP, NCOEF };

struct Batch {
    int nsys = 0;
    std::vector<SysGeom> geom;           // host copy
    DevBuf<SysGeom> geom_d;
    long ncells = 0;
    int ntiles = 0;
    int cpt = 2;                         // cells per thread; tile = BLOCK*cpt cells
    bool quad = false;                   // every system's nz is a multiple of 4: the
                                         // iteration kernels may take 4 cells of a z-row
                                         // per thread (cpt 4 and 8)
    int tiling_gen = 0;                  // bumped whenever tiles, grid or columns change
    int col_split = 0;                   // > 0: the mixed stencil marches columns, each cut
                                         // into this many r-segments (0: tile kernels)
    std::vector<ColDesc> cols_h;
    DevBuf<ColDesc> cols;
    std::vector<TileDesc> tiles_h;
    DevBuf<TileDesc> tiles;
    int nltiles = 0;                     // line tiles (0: the batch has no z-line kernels)
    DevBuf<LineTile> ltiles;
    DevBuf<double> coef[NCOEF];          // FP64 operand arenas, compact
    DevBuf<double> b;                    // Su plus the folded wall-halo terms of phi0
    DevBuf<double> x0;                   // initial phi interior
    DevBuf<double> fold;                 // wall-halo terms of phi0 (r = Su + fold - A_int*x)
    DevBuf<unsigned char> comp;          // projected-component slot per cell (NO_COMP: none)
    std::vector<CompInfo> comp_info;
    std::vector<DevBuf<double>*> phi_pad;  // padded phi per system (device)
    std::vector<char> uploaded;          // 0: no; 1: yes; 2: yes, singular components analysed
    DevBuf<double> staging;              // padded staging for uploads
    DevBuf<unsigned char> comp_bits;     // link and anchor bits of the component pre-pass
    DevBuf<int> comp_flag;
    hipStream_t stream = nullptr;
    int grid = 0;                        // blocks per tile-walking launch
    Cg64Work* cg64 = nullptr;
    MixedWork* mixed = nullptr;

    ~Batch();
    int init(int n, const eafp_desc* descs);
    int set_tiling(int cells_per_thread, int blocks);
    int set_columns(int r_splits);
    int upload_tables();
    // components = false skips the singular-component analysis, which only
    // the mixed solver needs (it reads the coefficients on the host).
    int upload(int s, const double* const host[8], const double* phi0, int arrays_on_device,
               bool components = true);
    int download(int s, double* phi_host, int to_device);
    int device_components(int s, const double* const A[7], int* singular);
    long padded_len(int s) const {
        const SysGeom& g = geom[s];
        return (long)(g.nr + 4) * (g.nth + 4) * (g.nz + 4);
    }
};

int validate_desc(const eafp_desc* d);
int pp_assemble(const eafp_desc* d, const eafp_pp_fields* f, const eafp_pp_config* cfg, double* const out[8],
                double* residual, double* flux_ref, const eafp_pp_exports* ex, int on_device);
int pp_correct(const eafp_desc* d, const eafp_pp_fields* f, const eafp_pp_config* cfg, double alpha_p,
               const double* pp, const eafp_pp_exports* ex, const eafp_pp_update* io, int on_device);
void pp_release();
int pp_step(const eafp_desc* d, const eafp_pp_fields* f, const eafp_pp_config* cfg, const eafp_pp_step_opts* o,
            double* pp, const eafp_pp_update* io, double* residual, eafp_stats* stats, int on_device);
void step_release();
int tdma_solve_batch(int nsys, const eafp_desc* descs, const eafp_tdma_system* sys, int n_sweep, int on_device);
void tdma_release();
int choose_grid(int ntiles);
int device_cus();
bool valid_cpt(int cpt);

}  // namespace eafp
