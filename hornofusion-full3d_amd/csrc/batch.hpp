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
#include <cstring>
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

// Instantiated graphs of one solver, kept until the solver's workspace is
// freed.  An entry's key is the solver's integer settings (format,
// preconditioner, iterations per graph, launch-shape generation) together
// with the kernel-argument struct the capture baked into its nodes, byte for
// byte, so a reallocated buffer is a new key and a graph never outlives the
// buffers it names.  Switching between keys never destroys a graph between
// solves; entries are dropped only once kMax are held, after the stream is
// idle, stale launch shapes first.  Args must be trivially copyable with
// deterministic padding (the solvers value-initialise theirs).
template <typename Args>
struct GraphCache {
    static constexpr int kKeys = 4;
    static constexpr size_t kMax = 12;
    struct Entry {
        int key[kKeys];
        int gen;
        Args args;
        hipGraphExec_t exec;
    };
    std::vector<Entry> v;

    hipGraphExec_t find(const int (&key)[kKeys], int gen, const Args& args) const {
        for (const Entry& e : v)
            if (e.gen == gen && memcmp(e.key, key, sizeof(e.key)) == 0 && memcmp(&e.args, &args, sizeof(Args)) == 0)
                return e.exec;
        return nullptr;
    }
    // Makes room for one more entry; the caller's stream must be idle.
    void make_room(int gen) {
        if (v.size() < kMax) return;
        for (size_t i = 0; i < v.size() && v.size() >= kMax;) {
            if (v[i].gen != gen) {
                (void)hipGraphExecDestroy(v[i].exec);
                v.erase(v.begin() + (long)i);
            } else {
                ++i;
            }
        }
        while (v.size() >= kMax) {
            (void)hipGraphExecDestroy(v.front().exec);
            v.erase(v.begin());
        }
    }
    void add(const int (&key)[kKeys], int gen, const Args& args, hipGraphExec_t exec) {
        Entry e;
        memcpy(e.key, key, sizeof(e.key));
        e.gen = gen;
        memcpy(&e.args, &args, sizeof(Args));
        e.exec = exec;
        v.push_back(e);
    }
    void clear() {
        for (Entry& e : v) (void)hipGraphExecDestroy(e.exec);
        v.clear();
    }
};

// Singular (pure-Neumann) connected components of one system, labelled on
// the device before its first mixed solve: the mixed solver projects its
// residuals and its result to zero mean on each of them.
struct CompInfo {
    int ncomp = 0;                 // projected components (<= MAX_COMP)
    int overflow = 0;              // more singular components than MAX_COMP
    int nsingular = 0;             // singular components found
    int cells[MAX_COMP] = {};      // cells of each projected component
    double inv_cnt[MAX_COMP] = {}; // 1 / cells of each
};

// Process-wide activity counters (eafp_get_counters).
enum { CTR_UPLOADS = 0, CTR_LABELLED, CTR_CAPTURES, CTR_LAUNCHES, CTR_COUNT };
void counter_add(int which, long long v);

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
    std::vector<char> uploaded;          // 0: no; 1: yes; 2: yes, singular components labelled
    DevBuf<double> staging;              // padded staging for uploads
    DevBuf<int> lab;                     // component labelling: a cell index per cell
    DevBuf<unsigned char> lab_bits;      // link, anchor and root bits per cell
    DevBuf<int> lab_small;               // per-system flags, root lists and counts
    long label_sweeps = 0;               // sweeps of the last labelling
    hipStream_t stream = nullptr;
    int grid = 0;                        // blocks per tile-walking launch
    Cg64Work* cg64 = nullptr;
    MixedWork* mixed = nullptr;

    ~Batch();
    int init(int n, const eafp_desc* descs);
    int set_tiling(int cells_per_thread, int blocks);
    int set_columns(int r_splits);
    int upload_tables();
    int upload(int s, const double* const host[8], const double* phi0, int arrays_on_device);
    int download(int s, double* phi_host, int to_device);
    // labels the singular components of the systems uploaded since their
    // last labelling (only the mixed solver needs them)
    int label_components();
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
int pp_assemble_noalloc(const eafp_desc* d, const eafp_pp_fields* f, const eafp_pp_config* cfg, double* const out[8],
                        double* residual, double* flux_ref, const eafp_pp_exports* ex, hipStream_t stream);
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
