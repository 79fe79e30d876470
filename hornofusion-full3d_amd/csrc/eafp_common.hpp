// eafp_common.hpp -- shared device/host definitions of the p' solver library.
//
// Layout on the device: every system is stored as its interior cells only,
// z fastest: cell (i, j, k), 0-based, sits at  off + (i*nth + j)*nz + k.
// The theta neighbour wraps by index arithmetic (the serial halo copy of
// src/mod_mpi_topology.f90:303-310 made implicit); r and z neighbours
// outside the mesh contribute a zero value, which is what the reference
// reads from the zeroed halos of its CG work vectors.
//
// Work decomposition: each system is cut into tiles of TILE_CELLS
// consecutive cells.  All tiles of all systems form one table; a launch
// walks the table with blocks pinned to XCDs (consecutive tiles run on the
// same XCD so stencil neighbours hit the same L2).  Per-tile partial sums
// are written to a partials array and the last block to finish a system's
// tiles reduces them in tile order (deterministic) and advances that
// system's scalar state.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace eafp {

constexpr int BLOCK = 256;                 // 4 wavefronts of 64
constexpr int CELLS_PER_THREAD = 2;
constexpr int TILE_CELLS = BLOCK * CELLS_PER_THREAD;
constexpr int NXCD = 8;                    // MI355X: 8 XCDs, blocks dealt round robin
constexpr double SMALL = 1.0e-30;          // src/mod_constants.f90:53
constexpr int MAX_COMP = 8;                // projected (singular) components per system

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

__device__ __forceinline__ CellIdx decompose(int local, const SysGeom& g) {
    int col = fdiv(local, g.nz, g.inv_nz);     // i*nth + j
    int k = local - col * g.nz;
    int i = fdiv(col, g.nth, g.inv_nth);
    int j = col - i * g.nth;
    return {i, j, k};
}

// ---------------------------------------------------------------------------
// Reductions.  Wave reduction by xor shuffles (64 lanes), block reduction
// through LDS.  The summation order is fixed, so every block that reduces
// the same data gets the same bits.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sum of v over the block, returned to every thread.  `sh` needs BLOCK/64
// entries per value.  Ends with a barrier so `sh` may be reused at once.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* sh) {
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < NV; ++c) {
        double w = wave_sum(v[c]);
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

// Reduce NV partial streams of `count` tiles (layout partial[t*stride + c])
// in tile order.  Called by one whole block; result to every thread.
template <int NV>
__device__ __forceinline__ void reduce_tiles(const double* partial, int stride, int t0, int t1,
                                             double (&out)[NV], double* sh) {
#pragma unroll
    for (int c = 0; c < NV; ++c) out[c] = 0.0;
    for (int t = t0 + (int)threadIdx.x; t < t1; t += BLOCK) {
#pragma unroll
        for (int c = 0; c < NV; ++c) out[c] += __builtin_nontemporal_load(&partial[(long)t * stride + c]);
    }
    block_sum<NV>(out, sh);
}

// Tile walk with XCD-contiguous segments: XCD x (= blockIdx % 8) handles
// tiles [ntiles*x/8, ntiles*(x+1)/8), its blocks striding through them.
struct TileWalk {
    int t, end, stride;
    __device__ TileWalk(int ntiles) {
        const int x = blockIdx.x % NXCD;
        const int slot = blockIdx.x / NXCD;
        const int nslots = gridDim.x / NXCD;
        const int b = (int)(((long)ntiles * x) / NXCD);
        end = (int)(((long)ntiles * (x + 1)) / NXCD);
        t = b + slot;
        stride = nslots;
    }
    __device__ bool valid() const { return t < end; }
    __device__ void next() { t += stride; }
};

// Last-block-done: every block that finished a tile of system s publishes
// its partial, then counts itself in; the block that brings the count to
// the number of tiles of s owns the reduction.  Returns true to every
// thread of that block only.
__device__ __forceinline__ bool arrive_last(unsigned* counter, unsigned ntiles, int* sh_flag) {
    __threadfence();
    if (threadIdx.x == 0) {
        unsigned prev = atomicAdd(counter, 1u);
        *sh_flag = (prev == ntiles - 1u);
    }
    __syncthreads();
    bool last = *sh_flag != 0;
    if (last) __threadfence();
    return last;
}

}  // namespace eafp
