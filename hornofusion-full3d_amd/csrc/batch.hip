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
// Device pre-pass of the singular-component analysis.  A component of the
// link graph is singular when every one of its rows is a Neumann row (the
// rule of analyze_components below); a linked row that is not a Neumann row
// anchors its component.  k_comp_init stores per compact cell the six link
// bits (a link exists when either side's coefficient is nonzero, theta
// periodic, no link across the r/z walls) and the anchor bit; k_comp_sweep
// spreads the anchor bit along links until a sweep changes nothing.  Then
// every linked cell is anchored exactly when the system has no singular
// component, and only otherwise does the host analysis run.
// ---------------------------------------------------------------------------
constexpr unsigned char LB_W = 1, LB_E = 2, LB_S = 4, LB_N = 8, LB_B = 16, LB_T = 32, LB_LINKS = 63, LB_ANCHOR = 64;

struct CompArgs {
    const double* A[7];   // padded aW..aT, aP (device)
    unsigned char* bits;  // per compact cell of the system
    int* flag;
};

__global__ void k_comp_init(CompArgs a, SysGeom g) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.n) return;
    const CellIdx id = decompose(c, g);
    const int i = id.i, j = id.j, k = id.k;
    const long sy = g.nr + 4, sz = (long)(g.nr + 4) * (g.nth + 4);
    const long o = (i + 2) + sy * (j + 2) + sz * (k + 2);
    const int jn = j + 1 == g.nth ? 0 : j + 1, jp = j == 0 ? g.nth - 1 : j - 1;
    const long on = o + sy * (jn - j), op = o + sy * (jp - j);
    const double *AW = a.A[C_AW], *AE = a.A[C_AE], *AS = a.A[C_AS], *AN = a.A[C_AN];
    const double *AB = a.A[C_AB], *AT = a.A[C_AT];
    unsigned char b = 0;
    if (i > 0 && (AE[o - 1] != 0.0 || AW[o] != 0.0)) b |= LB_W;
    if (i + 1 < g.nr && (AE[o] != 0.0 || AW[o + 1] != 0.0)) b |= LB_E;
    if (AN[op] != 0.0 || AS[o] != 0.0) b |= LB_S;
    if (AN[o] != 0.0 || AS[on] != 0.0) b |= LB_N;
    if (k > 0 && (AT[o - sz] != 0.0 || AB[o] != 0.0)) b |= LB_B;
    if (k + 1 < g.nz && (AT[o] != 0.0 || AB[o + sz] != 0.0)) b |= LB_T;
    // the Neumann test of analyze_components, in its order of additions
    double links = 0.0;
    if (i > 0) links += AW[o];
    if (i + 1 < g.nr) links += AE[o];
    links += AS[o] + AN[o];
    if (k > 0) links += AB[o];
    if (k + 1 < g.nz) links += AT[o];
    const double ap = a.A[C_AP][o];
    const bool wall = (i == 0 && AW[o] != 0.0) || (i + 1 == g.nr && AE[o] != 0.0) || (k == 0 && AB[o] != 0.0) ||
                      (k + 1 == g.nz && AT[o] != 0.0);
    const bool neumann = !wall && fabs(ap - links) <= 1e-10 * fabs(ap);
    if (b != 0 && !neumann) b |= LB_ANCHOR;
    a.bits[c] = b;
}

// One anchor-spreading sweep (in place: a sweep may also use what it set).
// flag: NULL, or set to 1 when the sweep anchors a cell.
__global__ void k_comp_sweep(unsigned char* bits, SysGeom g, int* flag) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.n) return;
    const unsigned char b = bits[c];
    if ((b & LB_LINKS) == 0 || (b & LB_ANCHOR)) return;
    const CellIdx id = decompose(c, g);
    const int plane = g.nth * g.nz;
    const int cs = id.j == 0 ? c + (g.nth - 1) * g.nz : c - g.nz;
    const int cn = id.j + 1 == g.nth ? c - (g.nth - 1) * g.nz : c + g.nz;
    bool hit = false;
    if (b & LB_W) hit |= (bits[c - plane] & LB_ANCHOR) != 0;
    if (b & LB_E) hit |= (bits[c + plane] & LB_ANCHOR) != 0;
    if (b & LB_S) hit |= (bits[cs] & LB_ANCHOR) != 0;
    if (b & LB_N) hit |= (bits[cn] & LB_ANCHOR) != 0;
    if (b & LB_B) hit |= (bits[c - 1] & LB_ANCHOR) != 0;
    if (b & LB_T) hit |= (bits[c + 1] & LB_ANCHOR) != 0;
    if (hit) {
        bits[c] = b | LB_ANCHOR;
        if (flag) *flag = 1;
    }
}

// flag = 1 when a linked cell is still unanchored (a singular component).
__global__ void k_comp_unanchored(const unsigned char* bits, SysGeom g, int* flag) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.n) return;
    const unsigned char b = bits[c];
    if ((b & LB_LINKS) != 0 && !(b & LB_ANCHOR)) *flag = 1;
}

__global__ void k_set_int(int* p, int v) { *p = v; }

__global__ void k_fill_u8(unsigned char* p, int n, unsigned char v) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) p[c] = v;
}

// Returns EAFP_OK with *singular = 0 when the system has no singular
// component (its slots are then set to NO_COMP), *singular = 1 otherwise.
int Batch::device_components(int s, const double* const A[7], int* singular) {
    const SysGeom& g = geom[s];
    EAFP_TRY_RC(comp_bits.alloc(std::max<size_t>(comp_bits.n, (size_t)g.n)));
    if (!comp_flag.p) EAFP_TRY_RC(comp_flag.alloc(1));
    CompArgs a;
    for (int q = 0; q < 7; ++q) a.A[q] = A[q];
    a.bits = comp_bits.p;
    a.flag = comp_flag.p;
    const int blocks = (g.n + 255) / 256;
    k_comp_init<<<blocks, 256, 0, stream>>>(a, g);
    EAFP_TRY(hipGetLastError());
    // Every sweep that changes something anchors at least one cell, so
    // g.n / SWEEPS + 1 rounds always reach the fixed point.
    constexpr int SWEEPS = 16;
    int changed = 1;
    for (long round = 0; changed && round <= g.n / SWEEPS + 1; ++round) {
        for (int q = 0; q < SWEEPS - 1; ++q) k_comp_sweep<<<blocks, 256, 0, stream>>>(comp_bits.p, g, nullptr);
        k_set_int<<<1, 1, 0, stream>>>(comp_flag.p, 0);
        k_comp_sweep<<<blocks, 256, 0, stream>>>(comp_bits.p, g, comp_flag.p);
        EAFP_TRY(hipGetLastError());
        EAFP_TRY(hipMemcpyAsync(&changed, comp_flag.p, sizeof(int), hipMemcpyDeviceToHost, stream));
        EAFP_TRY(hipStreamSynchronize(stream));
    }
    if (changed) return EAFP_ERR_STATE;   // unreachable: the bound above is exact
    int un = 0;
    k_set_int<<<1, 1, 0, stream>>>(comp_flag.p, 0);
    k_comp_unanchored<<<blocks, 256, 0, stream>>>(comp_bits.p, g, comp_flag.p);
    EAFP_TRY(hipGetLastError());
    EAFP_TRY(hipMemcpyAsync(&un, comp_flag.p, sizeof(int), hipMemcpyDeviceToHost, stream));
    EAFP_TRY(hipStreamSynchronize(stream));
    *singular = un;
    if (!un) {
        k_fill_u8<<<blocks, 256, 0, stream>>>(comp.p + g.off, g.n, NO_COMP);
        EAFP_TRY(hipGetLastError());
        comp_info[s] = CompInfo{};
    }
    return EAFP_OK;
}

int Batch::upload(int s, const double* const host[8], const double* phi0, int on_device, bool components) {
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

    if (!components) {
        EAFP_TRY(hipStreamSynchronize(stream));
        uploaded[s] = 1;
        return EAFP_OK;
    }
    int singular = 1;
    EAFP_TRY_RC(device_components(s, src, &singular));
    if (!singular) {
        EAFP_TRY(hipStreamSynchronize(stream));
        uploaded[s] = 2;
        return EAFP_OK;
    }
    // singular components (host): needs the coefficients on the host
    std::vector<double> hbuf;
    const double* hc[7];
    if (on_device) {
        hbuf.resize((size_t)np * 7);
        for (int q = 0; q < 7; ++q)
            EAFP_TRY(hipMemcpyAsync(hbuf.data() + (size_t)q * np, src[q], sizeof(double) * np, hipMemcpyDeviceToHost,
                                    stream));
        EAFP_TRY(hipStreamSynchronize(stream));
        for (int q = 0; q < 7; ++q) hc[q] = hbuf.data() + (size_t)q * np;
    } else {
        for (int q = 0; q < 7; ++q) hc[q] = host[q];
    }
    std::vector<unsigned char> slot(g.n);
    EAFP_TRY_RC(analyze_components(g, hc, slot.data(), comp_info[s]));
    EAFP_TRY(hipMemcpyAsync(comp.p + g.off, slot.data(), g.n, hipMemcpyHostToDevice, stream));
    EAFP_TRY(hipStreamSynchronize(stream));
    uploaded[s] = 2;
    return EAFP_OK;
}

// ---------------------------------------------------------------------------
// Connected components of the link graph (union-find over nonzero links,
// either side) and their singularity: a component is singular (constant
// null vector) when every row has aP equal to the sum of its interior
// links up to rounding, i.e. no compliance, acoustic, Dirichlet or wall
// term on its diagonal.  Rows without any link are never projected.
// ---------------------------------------------------------------------------
namespace {
int uf_find(std::vector<int>& p, int a) {
    while (p[a] != a) {
        p[a] = p[p[a]];
        a = p[a];
    }
    return a;
}
}  // namespace

int analyze_components(const SysGeom& g, const double* const A[7], unsigned char* slot, CompInfo& info) {
    const long sy = g.nr + 4, sz = (long)(g.nr + 4) * (g.nth + 4);
    auto po = [&](int i, int j, int k) { return (long)(i + 2) + sy * (j + 2) + sz * (k + 2); };
    auto cidx = [&](int i, int j, int k) { return (i * g.nth + j) * g.nz + k; };
    std::vector<int> parent(g.n);
    std::vector<char> linked(g.n, 0), neumann(g.n, 1);
    for (int c = 0; c < g.n; ++c) parent[c] = c;
    auto unite = [&](int a, int b) {
        a = uf_find(parent, a);
        b = uf_find(parent, b);
        if (a != b) parent[a < b ? b : a] = a < b ? a : b;
    };
    for (int i = 0; i < g.nr; ++i)
        for (int j = 0; j < g.nth; ++j)
            for (int k = 0; k < g.nz; ++k) {
                const long o = po(i, j, k);
                const int c = cidx(i, j, k);
                const int jn = (j + 1) % g.nth;
                double links = 0.0;
                if (i + 1 < g.nr && (A[C_AE][o] != 0.0 || A[C_AW][po(i + 1, j, k)] != 0.0)) {
                    unite(c, cidx(i + 1, j, k));
                    linked[c] = linked[cidx(i + 1, j, k)] = 1;
                }
                if (A[C_AN][o] != 0.0 || A[C_AS][po(i, jn, k)] != 0.0) {
                    unite(c, cidx(i, jn, k));
                    linked[c] = linked[cidx(i, jn, k)] = 1;
                }
                if (k + 1 < g.nz && (A[C_AT][o] != 0.0 || A[C_AB][po(i, j, k + 1)] != 0.0)) {
                    unite(c, cidx(i, j, k + 1));
                    linked[c] = linked[cidx(i, j, k + 1)] = 1;
                }
                if (i > 0) links += A[C_AW][o];
                if (i + 1 < g.nr) links += A[C_AE][o];
                links += A[C_AS][o] + A[C_AN][o];
                if (k > 0) links += A[C_AB][o];
                if (k + 1 < g.nz) links += A[C_AT][o];
                const double ap = A[C_AP][o];
                const bool wall = (i == 0 && A[C_AW][o] != 0.0) || (i + 1 == g.nr && A[C_AE][o] != 0.0) ||
                                  (k == 0 && A[C_AB][o] != 0.0) || (k + 1 == g.nz && A[C_AT][o] != 0.0);
                neumann[c] = !wall && std::fabs(ap - links) <= 1e-10 * std::fabs(ap);
            }
    // a component is singular iff all its rows are Neumann rows
    std::vector<char> root_sing(g.n, 1);
    for (int c = 0; c < g.n; ++c) {
        const int r = uf_find(parent, c);
        if (!linked[c] || !neumann[c]) root_sing[r] = 0;
    }
    std::vector<int> root_slot(g.n, -1);
    std::vector<long> cnt;
    info = CompInfo{};
    for (int c = 0; c < g.n; ++c) {
        slot[c] = NO_COMP;
        if (!linked[c]) continue;
        const int r = uf_find(parent, c);
        if (!root_sing[r]) continue;
        if (root_slot[r] < 0) {
            root_slot[r] = (int)cnt.size();
            cnt.push_back(0);
        }
        const int sl = root_slot[r];
        ++cnt[sl];
        if (sl < MAX_COMP) slot[c] = (unsigned char)sl;
    }
    info.ncomp = (int)std::min<size_t>(cnt.size(), MAX_COMP);
    info.overflow = cnt.size() > (size_t)MAX_COMP;
    for (int q = 0; q < info.ncomp; ++q) info.inv_cnt[q] = 1.0 / (double)cnt[q];
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
