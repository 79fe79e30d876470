// batch.hip -- system container: tile table, uploads and downloads between
// the reference's padded Fortran layout and the compact z-fastest layout.
#include <algorithm>
#include <cstring>

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

int choose_grid(int ntiles) {
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
            cus = v;
    }
    // Up to 8 blocks of 4 waves per CU; never more blocks than tiles.
    int g = std::min(ntiles, cus * 8);
    g = (g + NXCD - 1) / NXCD * NXCD;
    return std::max(g, NXCD);
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
    long off = 0;
    int t = 0;
    tile_sys_h.clear();
    tile_start_h.clear();
    for (int s = 0; s < n; ++s) {
        SysGeom& g = geom[s];
        g.nr = descs[s].nr;
        g.nth = descs[s].ntheta;
        g.nz = descs[s].nz;
        g.n = g.nr * g.nth * g.nz;
        g.off = off;
        g.inv_nz = 1.0f / (float)g.nz;
        g.inv_nth = 1.0f / (float)g.nth;
        g.tile_begin = t;
        for (int c = 0; c < g.n; c += TILE_CELLS) {
            tile_sys_h.push_back(s);
            tile_start_h.push_back(c);
            ++t;
        }
        g.tile_end = t;
        // keep every system's first cell 256-byte aligned for FP64 arrays
        off += (g.n + 31) / 32 * 32;
    }
    ncells = off;
    ntiles = t;
    EAFP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    EAFP_TRY_RC(geom_d.alloc(n));
    EAFP_TRY(hipMemcpy(geom_d.p, geom.data(), sizeof(SysGeom) * n, hipMemcpyHostToDevice));
    EAFP_TRY_RC(tile_sys.alloc(ntiles));
    EAFP_TRY_RC(tile_start.alloc(ntiles));
    EAFP_TRY(hipMemcpy(tile_sys.p, tile_sys_h.data(), sizeof(int) * ntiles, hipMemcpyHostToDevice));
    EAFP_TRY(hipMemcpy(tile_start.p, tile_start_h.data(), sizeof(int) * ntiles, hipMemcpyHostToDevice));
    for (int c = 0; c < NCOEF; ++c) EAFP_TRY_RC(coef[c].alloc(ncells));
    EAFP_TRY_RC(b.alloc(ncells));
    EAFP_TRY_RC(x0.alloc(ncells));
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
    grid = choose_grid(ntiles);
    return EAFP_OK;
}

// ---------------------------------------------------------------------------
// Pack: padded Fortran (r fastest) -> compact (z fastest).  One thread per
// compact cell; the wall links of phi0's r/z halos are folded into b:
//   b = Su + [i=1] aW*phi(0) + [i=nr] aE*phi(nr+1) + [k=1] aB*phi(.,.,0)
//          + [k=nz] aT*phi(.,.,nz+1)
// which is exactly the halo part of r = Su - A*phi at the first matvec
// (src/mod_solver_3d.f90:510-513); later CG vectors have zero r/z halos.
// ---------------------------------------------------------------------------
struct PackArgs {
    const double* src[8];   // padded aW..aP, Su (device staging)
    const double* phi;      // padded phi0 (device)
    double* dst[7];         // compact coefficient arenas (at system offset)
    double* b;
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
    double bb = a.src[7][o];
    const double* ph = a.phi;
    if (id.i == 0) bb += a.src[C_AW][o] * ph[o - 1];
    if (id.i == g.nr - 1) bb += a.src[C_AE][o] * ph[o + 1];
    if (id.k == 0) bb += a.src[C_AB][o] * ph[o - sz];
    if (id.k == g.nz - 1) bb += a.src[C_AT][o] * ph[o + sz];
    a.b[c] = bb;
    a.x0[c] = ph[o];
}

int Batch::upload(int s, const double* const host[8], const double* phi0, int on_device) {
    if (s < 0 || s >= nsys) return EAFP_ERR_STATE;
    for (int q = 0; q < 8; ++q)
        if (!host[q]) return EAFP_ERR_ARG;
    const long np = padded_len(s);
    EAFP_TRY_RC(staging.alloc(std::max<size_t>(staging.n, (size_t)np * 8)));
    const hipMemcpyKind kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    for (int q = 0; q < 8; ++q)
        EAFP_TRY(hipMemcpyAsync(staging.p + (long)q * np, host[q], sizeof(double) * np, kind, stream));
    if (phi0)
        EAFP_TRY(hipMemcpyAsync(phi_pad[s]->p, phi0, sizeof(double) * np, kind, stream));
    else
        EAFP_TRY(hipMemsetAsync(phi_pad[s]->p, 0, sizeof(double) * np, stream));
    const SysGeom& g = geom[s];
    PackArgs a;
    for (int q = 0; q < 8; ++q) a.src[q] = staging.p + (long)q * np;
    a.phi = phi_pad[s]->p;
    for (int q = 0; q < 7; ++q) a.dst[q] = coef[q].p + g.off;
    a.b = b.p + g.off;
    a.x0 = x0.p + g.off;
    k_pack<<<(g.n + 255) / 256, 256, 0, stream>>>(a, g);
    EAFP_TRY(hipGetLastError());
    EAFP_TRY(hipStreamSynchronize(stream));
    uploaded[s] = 1;
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
