// batch.hip -- system container: tile table, uploads and downloads between
// the reference's padded Fortran layout and the compact z-fastest layout.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "batch.hpp"

namespace eafp {

int validate_desc(const eafp_desc* d) {
    if (!d) return EAFP_ERR_ARG;
    if (d->is_parallel != 0) return EAFP_ERR_UNSUPPORTED;
    if (d->halo != 2 || d->lbound != -1) return EAFP_ERR_UNSUPPORTED;
    if (d->periodic_theta != 1) return EAFP_ERR_UNSUPPORTED;
    if (d->nr < 1 || d->ntheta < 1 || d->nz < 1) return EAFP_ERR_ARG;
    long n = (long)d->nr * d->ntheta * d->nz;
    if (n >= (1L << 24)) return EAFP_ERR_UNSUPPORTED;   // float-reciprocal index division
    return EAFP_OK;
}

int device_cus() {
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
            cus = v;
    }
    return cus;
}

int choose_grid(int ntiles) {
    // Up to 8 blocks of 4 waves per CU; never more blocks than tiles.
    int g = std::min(ntiles, device_cus() * 8);
    g = (g + NXCD - 1) / NXCD * NXCD;
    return std::max(g, NXCD);
}

bool valid_cpt(int cpt) { return cpt == 1 || cpt == 2 || cpt == 4 || cpt == 8; }

// Rebuild the tile table for `cells_per_thread` and the launch grid
// (`blocks` = 0: one block per tile up to 8 per CU).
int Batch::set_tiling(int cells_per_thread, int blocks) {
    if (!valid_cpt(cells_per_thread) || blocks < 0) return EAFP_ERR_ARG;
    const int tile = BLOCK * cells_per_thread;
    tiles_h = make_tile_table(geom, tile);
    ntiles = (int)tiles_h.size();
    cpt = cells_per_thread;
    EAFP_TRY_RC(tiles.alloc(ntiles));
    if (blocks > 0) {
        grid = std::max(NXCD, (blocks + NXCD - 1) / NXCD * NXCD);
    } else {
        grid = choose_grid(ntiles);
    }
    ++tiling_gen;
    // the column table refers to the tile ranges: rebuild it (or drop it
    // when the new tiles leave a system fewer slots than segments); the
    // tables go up once the stencil form is settled
    if (col_split > 0 && set_columns(col_split) == EAFP_OK) return EAFP_OK;
    col_split = 0;
    return upload_tables();
}

// Column-marching stencil for the mixed solver: r_splits > 0 cuts every
// column into that many r-segments; 0 returns to the tile kernels.
// EAFP_ERR_UNSUPPORTED when a system does not qualify (nz above
// COL_MAX_NZ, or fewer tiles than segments).
int Batch::set_columns(int r_splits) {
    if (r_splits < 0) return EAFP_ERR_ARG;
    if (r_splits == 0) {
        col_split = 0;
        ++tiling_gen;
        return upload_tables();
    }
    std::vector<ColDesc> t;
    if (!make_col_table(geom, r_splits, t)) return EAFP_ERR_UNSUPPORTED;
    cols_h = t;
    EAFP_TRY_RC(cols.alloc(cols_h.size()));
    EAFP_TRY(hipMemcpy(cols.p, cols_h.data(), sizeof(ColDesc) * cols_h.size(), hipMemcpyHostToDevice));
    col_split = r_splits;
    ++tiling_gen;
    return upload_tables();
}

// The tile table, the line tiles and the geometry on the device, with the
// p.q partial-slot counts of the current stencil form (col_split).
int Batch::upload_tables() {
    set_partial_counts(geom, tiles_h, col_split);
    EAFP_TRY(hipMemcpy(tiles.p, tiles_h.data(), sizeof(TileDesc) * ntiles, hipMemcpyHostToDevice));
    if (nltiles > 0) {   // the line tiles carry the tile ranges and counts: rebuild them too
        const std::vector<LineTile> lt = make_line_table(geom);
        EAFP_TRY_RC(ltiles.alloc(lt.size()));
        EAFP_TRY(hipMemcpy(ltiles.p, lt.data(), sizeof(LineTile) * lt.size(), hipMemcpyHostToDevice));
    }
    if (geom_d.p)
        EAFP_TRY(hipMemcpy(geom_d.p, geom.data(), sizeof(SysGeom) * nsys, hipMemcpyHostToDevice));
    return EAFP_OK;
}

void free_cg64(Cg64Work* w);
void free_mixed(MixedWork* w);

Batch::~Batch() {
    if (stream) (void)hipStreamSynchronize(stream);
    free_cg64(cg64);
    free_mixed(mixed);
    for (auto* p : phi_pad) delete p;
    if (stream) (void)hipStreamDestroy(stream);
}

int Batch::init(int n, const eafp_desc* descs) {
    if (n < 1 || !descs) return EAFP_ERR_ARG;
    for (int s = 0; s < n; ++s) EAFP_TRY_RC(validate_desc(&descs[s]));
    nsys = n;
    geom.resize(n);
    for (int s = 0; s < n; ++s) {
        geom[s].nr = descs[s].nr;
        geom[s].nth = descs[s].ntheta;
        geom[s].nz = descs[s].nz;
    }
    ncells = layout_systems(geom);
    nltiles = line_ranges(geom);
    quad = true;
    for (const SysGeom& g : geom) quad = quad && (g.nz % 4 == 0);
    EAFP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    EAFP_TRY_RC(geom_d.alloc(n));
    // default launch shape: for one system (the drop-in's case) what the
    // autotuner picks for the FP64 CG on the 60x120x84 production mesh, one
    // cell per thread and four blocks per CU (at most one per tile); for
    // batches, two cells per thread and one block per tile up to 8 per CU
    if (n == 1) {
        const int nt1 = (geom[0].n + BLOCK - 1) / BLOCK;
        EAFP_TRY_RC(set_tiling(1, std::min(4 * device_cus(), nt1)));
    } else {
        EAFP_TRY_RC(set_tiling(2, 0));
    }
    for (int c = 0; c < NCOEF; ++c) EAFP_TRY_RC(coef[c].alloc(ncells));
    EAFP_TRY_RC(b.alloc(ncells));
    EAFP_TRY_RC(x0.alloc(ncells));
    EAFP_TRY_RC(fold.alloc(ncells));
    EAFP_TRY_RC(comp.alloc(ncells));
    EAFP_TRY(hipMemset(fold.p, 0, sizeof(double) * ncells));
    EAFP_TRY(hipMemset(comp.p, NO_COMP, ncells));
    comp_info.assign(n, CompInfo{});
    EAFP_TRY(hipMemset(b.p, 0, sizeof(double) * ncells));
    EAFP_TRY(hipMemset(x0.p, 0, sizeof(double) * ncells));
    for (int c = 0; c < NCOEF; ++c) EAFP_TRY(hipMemset(coef[c].p, 0, sizeof(double) * ncells));
    phi_pad.resize(n, nullptr);
    for (int s = 0; s < n; ++s) {
        phi_pad[s] = new DevBuf<double>();
        EAFP_TRY_RC(phi_pad[s]->alloc(padded_len(s)));
        EAFP_TRY(hipMemset(phi_pad[s]->p, 0, sizeof(double) * padded_len(s)));
    }
    uploaded.assign(n, 0);
    return EAFP_OK;
}

// ---------------------------------------------------------------------------
// Pack: padded Fortran (r fastest) -> compact (z fastest).  One thread per
// compact cell.  The wall links of phi0's r/z halos are kept apart in
//   fold = [i=1] aW*phi(0) + [i=nr] aE*phi(nr+1) + [k=1] aB*phi(.,.,0)
//        + [k=nz] aT*phi(.,.,nz+1)
// which is exactly the halo part of r = Su - A*phi at the first matvec
// (src/mod_solver_3d.f90:510-513); later CG vectors have zero r/z halos.
// ---------------------------------------------------------------------------
struct PackArgs {
    const double* src[8];   // padded aW..aP, Su (device: staging or the caller's arrays)
    const double* phi;      // padded phi0 (device)
    double* dst[7];         // compact coefficient arenas (at system offset)
    double* b;
    double* fold;
    double* x0;
};

__global__ void k_pack(PackArgs a, SysGeom g) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.n) return;
    const CellIdx id = decompose(c, g);
    const long sy = g.nr + 4, sz = (long)(g.nr + 4) * (g.nth + 4);
    const long o = (id.i + 2) + sy * (id.j + 2) + sz * (id.k + 2);
#pragma unroll
    for (int q = 0; q < 7; ++q) a.dst[q][c] = a.src[q][o];
    const double* ph = a.phi;
    double f = 0.0;
    if (id.i == 0) f += a.src[C_AW][o] * ph[o - 1];
    if (id.i == g.nr - 1) f += a.src[C_AE][o] * ph[o + 1];
    if (id.k == 0) f += a.src[C_AB][o] * ph[o - sz];
    if (id.k == g.nz - 1) f += a.src[C_AT][o] * ph[o + sz];
    a.b[c] = a.src[7][o];
    a.fold[c] = f;
    a.x0[c] = ph[o];
}

// ---------------------------------------------------------------------------
// Singular components, labelled on the device.  A component of the link
// graph (a link exists where either side's coefficient is nonzero; theta
// periodic; none across the r/z walls) is singular -- its p' is fixed only
// up to a constant -- when every one of its rows is a Neumann row: aP equal
// to the sum of the row's interior links within 1e-10 relative, and no wall
// link.  The mixed solver projects its residuals and its result to zero
// mean on each singular component; rows without any link are never
// projected.  Kernels, one block per tile of the systems being labelled:
//   k_lab_init    link bits, the anchor bit (a linked row that is not a
//                 Neumann row), label = own cell index
//   k_lab_sweep   label = min over the linked neighbours' labels, then two
//                 pointer jumps (label = label of label: every label is the
//                 index of a cell of the same component, at most the
//                 labelled cell's own index); in place, until a sweep
//                 changes nothing, when every cell holds the smallest cell
//                 index of its component (its root)
//   k_lab_anchor  anchored cells mark their root
//   k_lab_roots   unmarked roots, the singular components, are listed
//   k_lab_assign  with each system's list sorted by the host (the order of
//                 first cells), a cell of a singular component takes its
//                 root's slot and is counted
// ---------------------------------------------------------------------------
constexpr unsigned char LB_W = 1, LB_E = 2, LB_S = 4, LB_N = 8, LB_B = 16, LB_T = 32, LB_LINKS = 63,
                        LB_ANCHOR = 64, LB_ROOT_ANCHORED = 128;
constexpr int ROOT_CAP = MAX_COMP + 1;

struct LabArgs {
    const TileDesc* tiles;
    int tile;                   // cells per tile
    const double* A[7];         // compact arenas aW..aT, aP
    unsigned char* bits;        // per arena cell
    int* lab;                   // per arena cell: a cell index of the same system
    const int* pending;         // per system: 1 = being labelled
    int* changed;
    int* nroots;                // per system
    int* roots;                 // per system: ROOT_CAP root indices (sorted by the host before k_lab_assign)
    int* counts;                // per system: MAX_COMP cell counts
    unsigned char* comp;        // projected-component slot per arena cell
};

struct LabCell {
    long o;                     // arena index
    int c;                      // cell index in the system
};

// The cells of block b's tile, when its system is being labelled.
#define LAB_FOR_CELLS(a, T, ...)                                                   \
    const TileDesc T = (a).tiles[blockIdx.x];                                       \
    if (!(a).pending[T.sys]) return;                                                \
    const int base_ = (blockIdx.x - T.tile_begin) * (a).tile;                       \
    const int end_ = min(T.n, base_ + (a).tile);                                    \
    for (int c = base_ + (int)threadIdx.x; c < end_; c += BLOCK) {                  \
        const long o = T.off + c;                                                   \
        __VA_ARGS__                                                                 \
    }

__global__ __launch_bounds__(BLOCK) void k_lab_init(LabArgs a) {
    LAB_FOR_CELLS(a, T, {
        const int k = c % T.nz, l = c / T.nz;
        const int j = l % T.nth, i = l / T.nth;
        const long plane = (long)T.nth * T.nz;
        const long oS = j == 0 ? o + (long)(T.nth - 1) * T.nz : o - T.nz;
        const long oN = j + 1 == T.nth ? o - (long)(T.nth - 1) * T.nz : o + T.nz;
        const double *AW = a.A[C_AW], *AE = a.A[C_AE], *AS = a.A[C_AS], *AN = a.A[C_AN];
        const double *AB = a.A[C_AB], *AT = a.A[C_AT];
        unsigned char b = 0;
        if (i > 0 && (AE[o - plane] != 0.0 || AW[o] != 0.0)) b |= LB_W;
        if (i + 1 < T.nr && (AE[o] != 0.0 || AW[o + plane] != 0.0)) b |= LB_E;
        if (AN[oS] != 0.0 || AS[o] != 0.0) b |= LB_S;
        if (AN[o] != 0.0 || AS[oN] != 0.0) b |= LB_N;
        if (k > 0 && (AT[o - 1] != 0.0 || AB[o] != 0.0)) b |= LB_B;
        if (k + 1 < T.nz && (AT[o] != 0.0 || AB[o + 1] != 0.0)) b |= LB_T;
        double links = 0.0;
        if (i > 0) links += AW[o];
        if (i + 1 < T.nr) links += AE[o];
        links += AS[o] + AN[o];
        if (k > 0) links += AB[o];
        if (k + 1 < T.nz) links += AT[o];
        const double ap = a.A[C_AP][o];
        const bool wall = (i == 0 && AW[o] != 0.0) || (i + 1 == T.nr && AE[o] != 0.0) || (k == 0 && AB[o] != 0.0) ||
                          (k + 1 == T.nz && AT[o] != 0.0);
        const bool neumann = !wall && fabs(ap - links) <= 1e-10 * fabs(ap);
        if (b != 0 && !neumann) b |= LB_ANCHOR;
        a.bits[o] = b;
        a.lab[o] = c;
    })
}

__global__ __launch_bounds__(BLOCK) void k_lab_sweep(LabArgs a, int flag) {
    LAB_FOR_CELLS(a, T, {
        const unsigned char b = a.bits[o];
        if ((b & LB_LINKS) == 0) continue;
        const int j = (c / T.nz) % T.nth;
        const long plane = (long)T.nth * T.nz;
        const long oS = j == 0 ? o + (long)(T.nth - 1) * T.nz : o - T.nz;
        const long oN = j + 1 == T.nth ? o - (long)(T.nth - 1) * T.nz : o + T.nz;
        const int cur = a.lab[o];
        int m = cur;
        if (b & LB_W) m = min(m, a.lab[o - plane]);
        if (b & LB_E) m = min(m, a.lab[o + plane]);
        if (b & LB_S) m = min(m, a.lab[oS]);
        if (b & LB_N) m = min(m, a.lab[oN]);
        if (b & LB_B) m = min(m, a.lab[o - 1]);
        if (b & LB_T) m = min(m, a.lab[o + 1]);
        m = a.lab[T.off + m];
        m = a.lab[T.off + m];
        if (m < cur) {
            a.lab[o] = m;
            if (flag) *a.changed = 1;
        }
    })
}

__global__ __launch_bounds__(BLOCK) void k_lab_anchor(LabArgs a) {
    LAB_FOR_CELLS(a, T, {
        const unsigned char b = a.bits[o];
        // only this bit of a root's byte changes in this kernel
        if ((b & LB_LINKS) && (b & LB_ANCHOR)) a.bits[T.off + a.lab[o]] |= LB_ROOT_ANCHORED;
    })
}

__global__ __launch_bounds__(BLOCK) void k_lab_roots(LabArgs a) {
    LAB_FOR_CELLS(a, T, {
        const unsigned char b = a.bits[o];
        if ((b & LB_LINKS) && !(b & LB_ROOT_ANCHORED) && a.lab[o] == c) {
            const int q = atomicAdd(&a.nroots[T.sys], 1);
            if (q < ROOT_CAP) a.roots[T.sys * ROOT_CAP + q] = c;
        }
    })
}

// The cell counts go to LDS first and leave the block as one global atomic
// per slot: a global atomic per cell serialises on the system's four
// counters (it took 129 ms for 64 systems of 604,800 cells).
__global__ __launch_bounds__(BLOCK) void k_lab_assign(LabArgs a) {
    __shared__ int cnt[MAX_COMP];
    const TileDesc T = a.tiles[blockIdx.x];
    if (!a.pending[T.sys]) return;
    if (threadIdx.x < MAX_COMP) cnt[threadIdx.x] = 0;
    __syncthreads();
    const int nr = a.nroots[T.sys];
    const int base_ = (blockIdx.x - T.tile_begin) * a.tile;
    const int end_ = min(T.n, base_ + a.tile);
    for (int c = base_ + (int)threadIdx.x; c < end_; c += BLOCK) {
        const long o = T.off + c;
        unsigned char slot = NO_COMP;
        if (nr <= MAX_COMP && (a.bits[o] & LB_LINKS)) {
            const int r = a.lab[o];
            if (!(a.bits[T.off + r] & LB_ROOT_ANCHORED)) {
                for (int q = 0; q < MAX_COMP; ++q)
                    if (q < nr && a.roots[T.sys * ROOT_CAP + q] == r) slot = (unsigned char)q;
                if (slot != NO_COMP) atomicAdd(&cnt[slot], 1);
            }
        }
        a.comp[o] = slot;
    }
    __syncthreads();
    if (threadIdx.x < MAX_COMP && cnt[threadIdx.x]) atomicAdd(&a.counts[T.sys * MAX_COMP + threadIdx.x], cnt[threadIdx.x]);
}

#undef LAB_FOR_CELLS

// Labels the singular components of every uploaded system not labelled since
// its upload (the mixed solver's need); fills comp and comp_info.
int Batch::label_components() {
    std::vector<int> pend(nsys, 0);
    int npend = 0;
    for (int s = 0; s < nsys; ++s)
        if (uploaded[s] == 1) {
            pend[s] = 1;
            ++npend;
        }
    if (npend == 0) return EAFP_OK;
    EAFP_TRY_RC(lab.alloc(ncells));
    EAFP_TRY_RC(lab_bits.alloc(ncells));
    EAFP_TRY_RC(lab_small.alloc((size_t)nsys * (2 + ROOT_CAP + MAX_COMP) + 1));
    int* pending_d = lab_small.p;
    int* nroots_d = pending_d + nsys;
    int* roots_d = nroots_d + nsys;
    int* counts_d = roots_d + (size_t)nsys * ROOT_CAP;
    int* changed_d = counts_d + (size_t)nsys * MAX_COMP;
    EAFP_TRY(hipMemcpyAsync(pending_d, pend.data(), sizeof(int) * nsys, hipMemcpyHostToDevice, stream));
    EAFP_TRY(hipMemsetAsync(nroots_d, 0, sizeof(int) * nsys, stream));
    EAFP_TRY(hipMemsetAsync(counts_d, 0, sizeof(int) * nsys * MAX_COMP, stream));
    LabArgs a;
    a.tiles = tiles.p;
    a.tile = BLOCK * cpt;
    for (int q = 0; q < 7; ++q) a.A[q] = coef[q].p;
    a.bits = lab_bits.p;
    a.lab = lab.p;
    a.pending = pending_d;
    a.changed = changed_d;
    a.nroots = nroots_d;
    a.roots = roots_d;
    a.counts = counts_d;
    a.comp = comp.p;
    k_lab_init<<<ntiles, BLOCK, 0, stream>>>(a);
    EAFP_TRY(hipGetLastError());
    // Labels only decrease, so the sweeps end; a check after every SWEEPS,
    // whose last sweep alone raises the flag: a sweep that changes nothing
    // is the fixed point.
    constexpr int SWEEPS = 4;
    int changed = 1;
    long sweeps = 0;
    while (changed) {
        if (sweeps > 4L * ncells + 64) return EAFP_ERR_STATE;   // unreachable: each changing sweep lowers a label
        for (int q = 0; q < SWEEPS - 1; ++q) k_lab_sweep<<<ntiles, BLOCK, 0, stream>>>(a, 0);
        EAFP_TRY(hipMemsetAsync(changed_d, 0, sizeof(int), stream));
        k_lab_sweep<<<ntiles, BLOCK, 0, stream>>>(a, 1);
        EAFP_TRY(hipGetLastError());
        EAFP_TRY(hipMemcpyAsync(&changed, changed_d, sizeof(int), hipMemcpyDeviceToHost, stream));
        EAFP_TRY(hipStreamSynchronize(stream));
        sweeps += SWEEPS;
    }
    label_sweeps = sweeps;
    k_lab_anchor<<<ntiles, BLOCK, 0, stream>>>(a);
    k_lab_roots<<<ntiles, BLOCK, 0, stream>>>(a);
    EAFP_TRY(hipGetLastError());
    std::vector<int> nroots(nsys), roots((size_t)nsys * ROOT_CAP);
    EAFP_TRY(hipMemcpyAsync(nroots.data(), nroots_d, sizeof(int) * nsys, hipMemcpyDeviceToHost, stream));
    EAFP_TRY(hipMemcpyAsync(roots.data(), roots_d, sizeof(int) * nsys * ROOT_CAP, hipMemcpyDeviceToHost, stream));
    EAFP_TRY(hipStreamSynchronize(stream));
    for (int s = 0; s < nsys; ++s)
        if (pend[s] && nroots[s] <= MAX_COMP) std::sort(roots.begin() + (long)s * ROOT_CAP,
                                                        roots.begin() + (long)s * ROOT_CAP + nroots[s]);
    EAFP_TRY(hipMemcpyAsync(roots_d, roots.data(), sizeof(int) * nsys * ROOT_CAP, hipMemcpyHostToDevice, stream));
    k_lab_assign<<<ntiles, BLOCK, 0, stream>>>(a);
    EAFP_TRY(hipGetLastError());
    std::vector<int> counts((size_t)nsys * MAX_COMP);
    EAFP_TRY(hipMemcpyAsync(counts.data(), counts_d, sizeof(int) * nsys * MAX_COMP, hipMemcpyDeviceToHost, stream));
    EAFP_TRY(hipStreamSynchronize(stream));
    for (int s = 0; s < nsys; ++s) {
        if (!pend[s]) continue;
        CompInfo& info = comp_info[s];
        info = CompInfo{};
        info.overflow = nroots[s] > MAX_COMP;
        info.ncomp = info.overflow ? 0 : nroots[s];
        for (int q = 0; q < info.ncomp; ++q) {
            info.cells[q] = counts[(size_t)s * MAX_COMP + q];
            info.inv_cnt[q] = 1.0 / (double)info.cells[q];
        }
        info.nsingular = nroots[s];
        uploaded[s] = 2;
    }
    counter_add(CTR_LABELLED, npend);
    return EAFP_OK;
}

int Batch::upload(int s, const double* const host[8], const double* phi0, int on_device) {
    if (s < 0 || s >= nsys) return EAFP_ERR_STATE;
    for (int q = 0; q < 8; ++q)
        if (!host[q]) return EAFP_ERR_ARG;
    const long np = padded_len(s);
    const hipMemcpyKind kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    // device arrays are packed where they lie; host arrays go through staging
    const double* src[8];
    if (on_device) {
        for (int q = 0; q < 8; ++q) src[q] = host[q];
    } else {
        EAFP_TRY_RC(staging.alloc(std::max<size_t>(staging.n, (size_t)np * 8)));
        for (int q = 0; q < 8; ++q) {
            EAFP_TRY(hipMemcpyAsync(staging.p + (long)q * np, host[q], sizeof(double) * np, kind, stream));
            src[q] = staging.p + (long)q * np;
        }
    }
    if (phi0)
        EAFP_TRY(hipMemcpyAsync(phi_pad[s]->p, phi0, sizeof(double) * np, kind, stream));
    else
        EAFP_TRY(hipMemsetAsync(phi_pad[s]->p, 0, sizeof(double) * np, stream));
    const SysGeom& g = geom[s];
    PackArgs a;
    for (int q = 0; q < 8; ++q) a.src[q] = src[q];
    a.phi = phi_pad[s]->p;
    for (int q = 0; q < 7; ++q) a.dst[q] = coef[q].p + g.off;
    a.b = b.p + g.off;
    a.fold = fold.p + g.off;
    a.x0 = x0.p + g.off;
    k_pack<<<(g.n + 255) / 256, 256, 0, stream>>>(a, g);
    EAFP_TRY(hipGetLastError());
    EAFP_TRY(hipStreamSynchronize(stream));
    uploaded[s] = 1;     // singular components: labelled by the first mixed solve that needs them
    counter_add(CTR_UPLOADS, 1);
    return EAFP_OK;
}

// ---------------------------------------------------------------------------
// Download: compact x -> interior of the padded phi copy, then the serial
// theta halo copy of src/mod_mpi_topology.f90:303-310, plane by plane in
// the reference's order, over every i and k including the r/z halos.
// ---------------------------------------------------------------------------
__global__ void k_unpack(const double* x, double* phi, SysGeom g) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.n) return;
    const CellIdx id = decompose(c, g);
    const long sy = g.nr + 4, sz = (long)(g.nr + 4) * (g.nth + 4);
    phi[(id.i + 2) + sy * (id.j + 2) + sz * (id.k + 2)] = x[c];
}

__global__ void k_theta_plane(double* phi, SysGeom g, int jdst, int jsrc) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int ni = g.nr + 4, nk = g.nz + 4;
    if (t >= ni * nk) return;
    const int i = t % ni, k = t / ni;
    const long sy = ni, sz = (long)ni * (g.nth + 4);
    phi[i + sy * (jdst + 1) + sz * k] = phi[i + sy * (jsrc + 1) + sz * k];
}

int download_x(Batch& B, int s, const double* x_compact) {
    const SysGeom& g = B.geom[s];
    double* phi = B.phi_pad[s]->p;
    k_unpack<<<(g.n + 255) / 256, 256, 0, B.stream>>>(x_compact + g.off, phi, g);
    const int nplane = (g.nr + 4) * (g.nz + 4);
    const int dst[4] = {0, -1, g.nth + 1, g.nth + 2};
    const int src[4] = {g.nth, g.nth - 1, 1, 2};
    for (int p = 0; p < 4; ++p)
        k_theta_plane<<<(nplane + 255) / 256, 256, 0, B.stream>>>(phi, g, dst[p], src[p]);
    EAFP_TRY(hipGetLastError());
    return EAFP_OK;
}

int Batch::download(int s, double* phi_out, int to_device) {
    if (s < 0 || s >= nsys) return EAFP_ERR_STATE;
    if (!phi_out) return EAFP_ERR_ARG;
    EAFP_TRY(hipMemcpyAsync(phi_out, phi_pad[s]->p, sizeof(double) * padded_len(s),
                            to_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, stream));
    EAFP_TRY(hipStreamSynchronize(stream));
    return EAFP_OK;
}

}  // namespace eafp
