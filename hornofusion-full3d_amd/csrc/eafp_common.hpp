// eafp_common.hpp -- shared device/host definitions of the p' solver library.
//
// Layout on the device: every system is stored as its interior cells only,
// z fastest: cell (i, j, k), 0-based, sits at  off + (i*nth + j)*nz + k.
// The theta neighbour wraps by index arithmetic (the serial halo copy of
// src/mod_mpi_topology.f90:303-310 made implicit); r and z neighbours
// outside the mesh contribute a zero value, which is what the reference
// reads from the zeroed halos of its CG work vectors.
//
// Work decomposition: each system is cut into tiles of BLOCK*CPT
// consecutive cells (CPT = cells per thread, chosen by the autotuner).  All
// tiles of all systems form one table; every block takes a contiguous
// range of it, with consecutive ranges on the same XCD so stencil
// neighbours hit the same L2.  A block accumulates each run of tiles of one
// system in registers and writes one partial sum at the run's first tile
// (zeros at the others); the next kernel reduces a system's partials in
// tile order at its start, so the reduction is deterministic and nothing
// crosses between blocks inside a kernel.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace eafp {

constexpr int BLOCK = 256;                 // 4 wavefronts of 64
constexpr int NXCD = 8;                    // MI355X: 8 XCDs, blocks dealt round robin
constexpr double SMALL = 1.0e-30;          // src/mod_constants.f90:53
constexpr int MAX_COMP = 4;                // projected (singular) components per system
constexpr unsigned char NO_COMP = 255;      // cell outside every projected component

// Per-tile descriptor: the tile's system and that system's geometry in
// one 48-byte record, so a block needs a single load to address its cells.
struct TileDesc {
    int sys;
    int tile_begin, tile_end;   // the system's tiles
    int n;                      // cells of the system
    int nr, nth, nz, pad;
    long off;                   // first cell of the system in the arenas
    float inv_nz, inv_nth;
};

// Geometry of one system (device-visible, read-only during a solve).
struct SysGeom {
    int nr, nth, nz;
    int n;            // nr*nth*nz
    long off;         // first cell of this system in the cell arenas
    int tile_begin;   // [tile_begin, tile_end) in the tile table
    int tile_end;
    float inv_nz;     // 1/nz, 1/nth for the float-reciprocal division
    float inv_nth;
};

// Quotient of u by d for u < 2^24 using a float reciprocal and a one-step
// integer correction (exact for every u < 2^24, d >= 1).
__device__ __forceinline__ int fdiv(int u, int d, float inv_d) {
    int q = (int)((float)u * inv_d);
    int r = u - q * d;
    if (r >= d) q += 1;
    else if (r < 0) q -= 1;
    return q;
}

struct CellIdx {
    int i, j, k;
};

// Offsets of the six stencil neighbours of compact cell o = off + c.  The
// theta pair wraps; an r/z neighbour beyond a wall is replaced by o itself
// (always a valid address) and flagged off, so every load can be issued
// unconditionally and the wall applied afterwards by a select: the loads of
// a cell never wait on a branch.
struct Nb {
    long W, E, S, N, B, T;
    bool hW, hE, hB, hT;
};

template <typename G>
__device__ __forceinline__ Nb neighbours(const CellIdx& id, const G& g, long o) {
    Nb n;
    const long sr = (long)g.nth * g.nz;
    n.hW = id.i > 0;
    n.hE = id.i < g.nr - 1;
    n.hB = id.k > 0;
    n.hT = id.k < g.nz - 1;
    n.W = n.hW ? o - sr : o;
    n.E = n.hE ? o + sr : o;
    n.S = id.j > 0 ? o - g.nz : o + (long)(g.nth - 1) * g.nz;
    n.N = id.j < g.nth - 1 ? o + g.nz : o - (long)(g.nth - 1) * g.nz;
    n.B = n.hB ? o - 1 : o;
    n.T = n.hT ? o + 1 : o;
    return n;
}

template <typename G>
__device__ __forceinline__ CellIdx decompose(int local, const G& g) {
    int col = fdiv(local, g.nz, g.inv_nz);     // i*nth + j
    int k = local - col * g.nz;
    int i = fdiv(col, g.nth, g.inv_nth);
    int j = col - i * g.nth;
    return {i, j, k};
}

// ---------------------------------------------------------------------------
// Reductions.  Within a wavefront by DPP (quad permutes, then row rotates
// by 4 and 8: every row of 16 lanes ends with its sum), the four row sums
// read from lanes 0, 16, 32, 48; across the four wavefronts through LDS.
// The order is fixed, so every block that reduces the same data gets the
// same bits.
// ---------------------------------------------------------------------------
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double readlane_d(double v, int lane) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

// Sum over the wavefront, returned uniformly.
__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_d<0xB1>(v);    // quad_perm [1,0,3,2]
    v += dpp_d<0x4E>(v);    // quad_perm [2,3,0,1]
    v += dpp_d<0x124>(v);   // row_ror:4
    v += dpp_d<0x128>(v);   // row_ror:8
    return (readlane_d(v, 0) + readlane_d(v, 16)) + (readlane_d(v, 32) + readlane_d(v, 48));
}

// Sum of v over the block, returned to every thread.  `sh` needs BLOCK/64
// entries per value.  Ends with a barrier so `sh` may be reused at once.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* sh) {
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < NV; ++c) {
        const double w = wave_sum(v[c]);
        if (lane == 0) sh[c * (BLOCK / 64) + wid] = w;
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < NV; ++c) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < BLOCK / 64; ++w) s += sh[c * (BLOCK / 64) + w];
        v[c] = s;
    }
    __syncthreads();
}

// Writes the block's run sum v (NV columns) at tile t0 and zeros at the
// other tiles of the run [t0, t1).
template <int NV>
__device__ __forceinline__ void store_run(double* partial, int t0, int t1, const double (&v)[NV]) {
    if (threadIdx.x == 0) {
#pragma unroll
        for (int c = 0; c < NV; ++c) partial[(long)t0 * NV + c] = v[c];
    }
    for (int u = t0 + 1 + (int)threadIdx.x; u < t1; u += BLOCK) {
#pragma unroll
        for (int c = 0; c < NV; ++c) partial[(long)u * NV + c] = 0.0;
    }
}

// Sum the NV-column partials of tiles [t0, t1) in tile order.  Called by
// a whole block; the result reaches every thread, and every block that
// reduces the same tiles gets the same bits.
template <int NV>
__device__ __forceinline__ void reduce_tiles(const double* partial, int t0, int t1, double (&out)[NV], double* sh) {
#pragma unroll
    for (int c = 0; c < NV; ++c) out[c] = 0.0;
    for (int t = t0 + (int)threadIdx.x; t < t1; t += BLOCK) {
#pragma unroll
        for (int c = 0; c < NV; ++c) out[c] += partial[(long)t * NV + c];
    }
    block_sum<NV>(out, sh);
}

// Contiguous tile range of this block.  Logical block lb = x*(G/8) + b/8
// for the block b on XCD x = b % 8 (blocks are dealt to XCDs round robin),
// so consecutive tiles -- stencil neighbours -- run on the same XCD and
// share its L2; a block's tiles touch few systems.
struct TileRange {
    int t, end;
    __device__ TileRange(int ntiles) {
        const int G = gridDim.x;
        const int lb = (blockIdx.x % NXCD) * (G / NXCD) + blockIdx.x / NXCD;
        t = (int)(((long)ntiles * lb) / G);
        end = (int)(((long)ntiles * (lb + 1)) / G);
    }
};

// Per-system scalar state lives in two slots.  Kernels come in two roles
// that alternate in every launch sequence: role 0 reads slot 1 and writes
// slot 0, role 1 reads slot 0 and writes slot 1.  A kernel never writes
// what it reads, so no block can see a half-updated state; the block that
// owns a system's first tile is the one that writes the system's new slot
// (every block derives the same values from the same partial sums).
// Nothing crosses between blocks inside a kernel: no atomics, no fences.

}  // namespace eafp
