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
// tiles of all systems form one table, cut into one contiguous chunk per
// XCD; the blocks of an XCD sweep its chunk together, block by block and
// tile by tile (TileWalk), so the r-neighbour planes a stencil reads were
// brought into that XCD's L2 by the other blocks of the same sweep.  A
// block accumulates the tiles it takes of one system (a run) in registers
// and writes one partial sum at the run's first tile (zeros at its others);
// the next kernel reduces a system's partials in tile order at its start,
// so the reduction is deterministic and nothing crosses between blocks
// inside a kernel.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <vector>

namespace eafp {

constexpr int BLOCK = 256;                 // 4 wavefronts of 64
constexpr int NXCD = 8;                    // MI355X: 8 XCDs, blocks dealt round robin
constexpr double SMALL = 1.0e-30;          // src/mod_constants.f90:53
constexpr int MAX_COMP = 4;                // projected (singular) components per system
constexpr unsigned char NO_COMP = 255;      // cell outside every projected component

// Per-tile descriptor: the tile's system and that system's geometry in
// one 56-byte record, so a block needs a single load to address its cells.
struct TileDesc {
    int sys;
    int tile_begin, tile_end;   // the system's tiles
    int n;                      // cells of the system
    int nr, nth, nz;
    int lt_begin, lt_end;       // the system's line tiles (z-line preconditioned kernels)
    int pa_n;                   // p.q partial slots the stencil writes: tile_begin .. tile_begin+pa_n-1
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
    int lt_begin;     // [lt_begin, lt_end) in the line-tile table (0, 0: none)
    int lt_end;
    int pa_n;         // p.q partial slots of the current stencil form (tiles, or column segments)
};

// Lays out the systems' cells in one arena: fills nr/nth/nz/n/off and the
// division reciprocals of every entry of g (whose nr, nth, nz are set) and
// returns the arena length.  Every system starts at a 32-cell boundary, so
// its FP64 arrays start 256-byte aligned.
inline long layout_systems(std::vector<SysGeom>& g) {
    long off = 0;
    for (SysGeom& s : g) {
        s.n = s.nr * s.nth * s.nz;
        s.off = off;
        s.inv_nz = 1.0f / (float)s.nz;
        s.inv_nth = 1.0f / (float)s.nth;
        off += (s.n + 31) / 32 * 32;
    }
    return off;
}

// The tile table for tiles of `tile` cells: each system's cells cut into
// consecutive tiles, systems in order.  Sets every system's tile range.
inline std::vector<TileDesc> make_tile_table(std::vector<SysGeom>& g, int tile) {
    std::vector<TileDesc> tiles;
    int t = 0;
    for (int s = 0; s < (int)g.size(); ++s) {
        SysGeom& G = g[s];
        const int nt = (G.n + tile - 1) / tile;
        G.tile_begin = t;
        G.tile_end = t + nt;
        TileDesc d{};
        d.sys = s;
        d.tile_begin = G.tile_begin;
        d.tile_end = G.tile_end;
        d.n = G.n;
        d.nr = G.nr;
        d.nth = G.nth;
        d.nz = G.nz;
        d.off = G.off;
        d.inv_nz = G.inv_nz;
        d.inv_nth = G.inv_nth;
        d.lt_begin = G.lt_begin;
        d.lt_end = G.lt_end;
        G.pa_n = nt;
        d.pa_n = nt;
        tiles.insert(tiles.end(), nt, d);
        t += nt;
    }
    return tiles;
}

// ---------------------------------------------------------------------------
// Line tiles, the work unit of the z-line preconditioned kernels of the
// mixed solver.  Line l = i*nth + j of a system is its z-row, the nz
// consecutive cells l*nz .. l*nz + nz-1.  A block takes lpb = BLOCK/nq whole
// lines of one system, nq = ceil(nz/4) lanes per line; lane j of line li
// holds the line's cells 4j .. 4j+3 (those below nz), so the block's lines
// are one contiguous run of cells.  When every nz of the batch is a multiple
// of 4, thread t holds the cells line0*nz + 4t .. +3 and reads them with
// 16-byte accesses.  Needs every nz at most 4*BLOCK.
// ---------------------------------------------------------------------------
struct LineTile {
    int sys;
    int line0, nlines;          // lines line0 .. line0+nlines-1 of the system
    int nz;
    int tile_begin, tile_end;   // the system's tiles (the stencil's partials)
    int lt_begin, lt_end;       // the system's line tiles (this kernel's partials)
    int pa_n;                   // the stencil's p.q partial slots (as TileDesc::pa_n)
    int pad;
    long off;                   // first cell of the system in the arenas
};

__host__ __device__ __forceinline__ int line_lanes(int nz) { return (nz + 3) / 4; }
__host__ __device__ __forceinline__ int line_lpb(int nz) { return BLOCK / line_lanes(nz); }

inline bool line_eligible(const std::vector<SysGeom>& g) {
    for (const SysGeom& s : g)
        if (s.nz < 1 || s.nz > 4 * BLOCK) return false;
    return true;
}

// Sets every system's line-tile range (consecutive, systems in order), or
// clears them all when the batch is not eligible.  Returns the tile count.
inline int line_ranges(std::vector<SysGeom>& g) {
    const bool ok = line_eligible(g);
    int t = 0;
    for (SysGeom& s : g) {
        const int nlt = ok ? (s.nr * s.nth + line_lpb(s.nz) - 1) / line_lpb(s.nz) : 0;
        s.lt_begin = t;
        s.lt_end = t + nlt;
        t += nlt;
    }
    return t;
}

// The line-tile table (after make_tile_table has set the tile ranges).
inline std::vector<LineTile> make_line_table(const std::vector<SysGeom>& g) {
    std::vector<LineTile> out;
    for (int s = 0; s < (int)g.size(); ++s) {
        const SysGeom& G = g[s];
        const int lines = G.nr * G.nth;
        const int lpb = G.lt_end > G.lt_begin ? line_lpb(G.nz) : 0;
        for (int t = G.lt_begin; t < G.lt_end; ++t) {
            LineTile L{};
            L.sys = s;
            L.line0 = (t - G.lt_begin) * lpb;
            L.nlines = lines - L.line0 < lpb ? lines - L.line0 : lpb;
            L.nz = G.nz;
            L.tile_begin = G.tile_begin;
            L.tile_end = G.tile_end;
            L.lt_begin = G.lt_begin;
            L.lt_end = G.lt_end;
            L.pa_n = G.pa_n;
            L.off = G.off;
            out.push_back(L);
        }
    }
    return out;
}

// ---------------------------------------------------------------------------
// Column segments, the work unit of the column-marching stencil kernels
// (systems with nz at most COL_MAX_NZ).  A column is a slab of nj
// theta-rows of one system (nj*nz <= 4*BLOCK cells: four cells per thread,
// one quad of a row when nz is a multiple of 4), marched through the
// r-planes i0 .. i1-1 of its segment.  The segments of a system are numbered 0 .. nseg-1; segment u
// writes its partial sum at tile slot tile_begin + u and zeros at the slots
// tile_begin + u + v*nseg (v >= 1) of the system, so the system's partials
// still cover its whole tile range exactly once.
// ---------------------------------------------------------------------------
constexpr int COL_MAX_NZ = 256;
constexpr int COL_CELLS = 4 * BLOCK;   // cells of one plane of a column, at most

struct ColDesc {
    int tile;      // the system's first tile; -1: empty
    int j0, nj;    // theta rows j0 .. j0+nj-1
    int i0, i1;    // r planes of the segment
    int seg;       // index among the system's segments (0 writes the system's state)
    int nseg;      // segments of the system
    int dir;       // +1: the march goes up in r, -1: down (odd segments of a column)
    // the system, as its TileDesc carries it: a block needs this one load
    // before it can address its cells
    int sys, nr, nth, nz;
    int tile_end;            // the system's tiles: tile .. tile_end-1
    int lt_begin, lt_end;    // the system's line tiles
    long off;                // first cell of the system in the arenas
};

// Rows per column for a given nz (every thread holds one quad of a row).
__host__ __device__ __forceinline__ int col_rows(int nz) { return COL_CELLS / nz; }

// Builds the column table: every system cut into ceil(nth/rows) columns and
// each column into `splits` r-segments (fewer when nr is smaller).  Padded
// with empty entries to a multiple of NXCD.  Returns false (and no table)
// when a system does not qualify: nz above COL_MAX_NZ, or fewer tiles than
// segments (the partial slots).
// Segments of one system at `splits` r-segments per column.
inline int col_segments(const SysGeom& G, int splits) {
    const int rows = col_rows(G.nz);
    return ((G.nth + rows - 1) / rows) * (splits < G.nr ? splits : G.nr);
}

// The p.q partial slots each system's stencil writes: one per column
// segment in the column form (the segment's index), one per tile otherwise.
// Sets SysGeom::pa_n and every tile's copy of it.
inline void set_partial_counts(std::vector<SysGeom>& g, std::vector<TileDesc>& tiles, int col_splits) {
    for (SysGeom& G : g) {
        const int nt = G.tile_end - G.tile_begin;
        const int ns = col_splits > 0 ? col_segments(G, col_splits) : nt;
        G.pa_n = ns < nt ? ns : nt;   // the column form exists only when every system has nseg <= its tiles
    }
    for (TileDesc& T : tiles) T.pa_n = g[T.sys].pa_n;
}

inline bool make_col_table(const std::vector<SysGeom>& g, int splits, std::vector<ColDesc>& out) {
    out.clear();
    for (const SysGeom& G : g) {
        if (G.nz > COL_MAX_NZ) return false;
        const int rows = col_rows(G.nz);
        const int ncol = (G.nth + rows - 1) / rows;
        const int nsp = splits < G.nr ? splits : G.nr;
        const int nseg = ncol * nsp;
        if (G.tile_end - G.tile_begin < nseg) return false;
        for (int c = 0; c < ncol; ++c)
            for (int sp = 0; sp < nsp; ++sp) {
                ColDesc d{};
                d.tile = G.tile_begin;
                d.sys = (int)(&G - g.data());
                d.nr = G.nr;
                d.nth = G.nth;
                d.nz = G.nz;
                d.tile_end = G.tile_end;
                d.lt_begin = G.lt_begin;
                d.lt_end = G.lt_end;
                d.off = G.off;
                d.j0 = c * rows;
                d.nj = (G.nth - d.j0) < rows ? (G.nth - d.j0) : rows;
                d.i0 = (int)(((long)G.nr * sp) / nsp);
                d.i1 = (int)(((long)G.nr * (sp + 1)) / nsp);
                d.seg = c * nsp + sp;
                d.nseg = nseg;
                d.dir = (sp & 1) ? -1 : 1;
                out.push_back(d);
            }
    }
    while (out.size() % NXCD) {
        ColDesc e{};
        e.tile = -1;
        e.dir = 1;
        out.push_back(e);
    }
    return true;
}

// Quotient of u by d for u < 2^24 using a float reciprocal and a one-step
// integer correction (exact for every u < 2^24, d >= 1).
__host__ __device__ __forceinline__ int fdiv(int u, int d, float inv_d) {
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
__host__ __device__ __forceinline__ Nb neighbours(const CellIdx& id, const G& g, long o) {
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

// A quad: 4 consecutive cells (k0 .. k0+3) of one z-row, for systems whose
// nz is a multiple of 4 (k0 is then one too).  Its W/E/S/N neighbours are
// the quads at the same k0 in the adjacent rows -- 16-byte aligned like the
// quad itself, since every system starts at a 32-cell boundary -- B is the
// cell below the first and T the cell above the last.  As in neighbours(),
// a wall neighbour is replaced by the quad's own first cell and flagged.
struct Nb4 {
    long W, E, S, N, B, T;
    bool hW, hE, hB, hT;
};

template <typename G>
__host__ __device__ __forceinline__ Nb4 neighbours4(const CellIdx& id, const G& g, long o) {
    Nb4 n;
    const long sr = (long)g.nth * g.nz;
    n.hW = id.i > 0;
    n.hE = id.i < g.nr - 1;
    n.hB = id.k > 0;
    n.hT = id.k + 4 < g.nz;
    n.W = n.hW ? o - sr : o;
    n.E = n.hE ? o + sr : o;
    n.S = id.j > 0 ? o - g.nz : o + (long)(g.nth - 1) * g.nz;
    n.N = id.j < g.nth - 1 ? o + g.nz : o - (long)(g.nth - 1) * g.nz;
    n.B = n.hB ? o - 1 : o;
    n.T = n.hT ? o + 4 : o;
    return n;
}

// Four floats moved as one 16-byte access (p 16-byte aligned).
struct F4 {
    float v[4];
};

__device__ __forceinline__ F4 ld4(const float* p) {
    F4 r;
    __builtin_memcpy(&r, __builtin_assume_aligned(p, 16), 16);
    return r;
}

__device__ __forceinline__ void st4(float* p, const F4& r) {
    __builtin_memcpy(__builtin_assume_aligned(p, 16), &r, 16);
}

// The same with the nontemporal hint, for streams that the kernel touches
// once: measured on MI355X, the update kernel's five loads and two of its
// stores went from 5.4 to 5.8 TB/s with it.  A kernel that uses these on
// an array accesses that array through no other type.
typedef float f4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ F4 ld4nt(const float* p) {
    const f4v v = __builtin_nontemporal_load(reinterpret_cast<const f4v*>(p));
    F4 r;
    r.v[0] = v.x;
    r.v[1] = v.y;
    r.v[2] = v.z;
    r.v[3] = v.w;
    return r;
}

__device__ __forceinline__ void st4nt(float* p, const F4& r) {
    const f4v v = {r.v[0], r.v[1], r.v[2], r.v[3]};
    __builtin_nontemporal_store(v, reinterpret_cast<f4v*>(p));
}

template <typename G>
__host__ __device__ __forceinline__ CellIdx decompose(int local, const G& g) {
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

// Writes the block's run sum v (NV columns) at the run's first tile t0 and
// zeros at its other n-1 tiles t0 + u*stride.
template <int NV>
__device__ __forceinline__ void store_run(double* partial, int t0, int n, int stride, const double (&v)[NV]) {
    if (threadIdx.x == 0) {
#pragma unroll
        for (int c = 0; c < NV; ++c) partial[(long)t0 * NV + c] = v[c];
    }
    for (int u = 1 + (int)threadIdx.x; u < n; u += BLOCK) {
        const long t = t0 + (long)u * stride;
#pragma unroll
        for (int c = 0; c < NV; ++c) partial[t * NV + c] = 0.0;
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

// reduce_tiles in two steps, so that a kernel can issue the partials' loads
// ahead of its own cell loads and add them after: each thread's first
// RT_REG partials are loaded into registers, then added with the rest in the
// order reduce_tiles adds them (the same bits).
constexpr int RT_REG = 4;
template <int NV>
struct TilePre {
    double v[RT_REG][NV];
};
template <int NV>
__device__ __forceinline__ void reduce_tiles_load(const double* partial, int t0, int t1, TilePre<NV>& pre) {
#pragma unroll
    for (int k = 0; k < RT_REG; ++k) {
        const int t = t0 + (int)threadIdx.x + k * BLOCK;
#pragma unroll
        for (int c = 0; c < NV; ++c) pre.v[k][c] = t < t1 ? partial[(long)t * NV + c] : 0.0;
    }
}
template <int NV>
__device__ __forceinline__ void reduce_tiles_sum(const double* partial, int t0, int t1, const TilePre<NV>& pre,
                                                 double (&out)[NV], double* sh) {
#pragma unroll
    for (int c = 0; c < NV; ++c) out[c] = 0.0;
#pragma unroll
    for (int k = 0; k < RT_REG; ++k)
        if (t0 + (int)threadIdx.x + k * BLOCK < t1) {
#pragma unroll
            for (int c = 0; c < NV; ++c) out[c] += pre.v[k][c];
        }
    for (int t = t0 + (int)threadIdx.x + RT_REG * BLOCK; t < t1; t += BLOCK) {
#pragma unroll
        for (int c = 0; c < NV; ++c) out[c] += partial[(long)t * NV + c];
    }
    block_sum<NV>(out, sh);
}

// The tiles of block b of a grid of G blocks (G a multiple of NXCD).  The
// hardware deals blocks to the XCDs round robin, block b to XCD x = b % 8;
// the table is cut into 8 contiguous chunks, chunk x for XCD x, and inside
// it the block takes tiles chunk_begin + b/8 + u*(G/8), u = 0, 1, ...  At
// any moment the blocks of one XCD thus work on one contiguous window of
// the table: a stencil's r-neighbour plane (nth*nz cells away) lies in the
// same window or the one just before it, already in that XCD's L2.  When G
// is at least the tile count every block takes at most one tile.
struct TileWalk {
    int first;    // the block's first tile (>= end: no tiles)
    int end;      // end of the XCD's chunk
    int stride;   // G / 8
    __host__ __device__ TileWalk(int ntiles, int G, int b) {
        const int x = b % NXCD;
        const int begin = (int)(((long)ntiles * x) / NXCD);
        end = (int)(((long)ntiles * (x + 1)) / NXCD);
        stride = G / NXCD;
        first = begin + b / NXCD;
    }
    // Tiles of the run that starts at the block's tile t and ends with the
    // system's last tile or the chunk: t, t + stride, ... < min(end, tile_end).
    __host__ __device__ int run(int t, int tile_end) const {
        const int e = tile_end < end ? tile_end : end;
        return (e - t + stride - 1) / stride;
    }
};

__device__ __forceinline__ TileWalk my_tiles(int ntiles) { return TileWalk(ntiles, (int)gridDim.x, (int)blockIdx.x); }

// Compact cell index of a tile's slot: the cell itself when it exists, else
// the system's last cell, so that a load for a slot past the end of a
// system's last (partial) tile stays inside the system.
__host__ __device__ __forceinline__ int clamp_cell(int c, int n) { return c < n ? c : n - 1; }

// Writes a small per-system state record with 8-byte nontemporal stores.
// The record is written by one thread of one block per kernel and read by
// every block of the next kernel; measured on MI355X, an ordinary store of
// it held the writing block back by ~13 us (a 604,800-cell update kernel
// went from 9 us to 22 us), a nontemporal store costs nothing measurable.
// The words are taken out of v with memcpy: reading the record through an
// integer pointer would break strict aliasing, and the compiler then may
// (and did) store words loaded before v's double fields were updated.
template <typename R>
__device__ __forceinline__ void store_state(R* dst, const R& v) {
    static_assert(sizeof(R) % 8 == 0, "state records are whole 8-byte words");
    unsigned long long* d = reinterpret_cast<unsigned long long*>(dst);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(R) / 8); ++i) {
        unsigned long long w;
        __builtin_memcpy(&w, reinterpret_cast<const char*>(&v) + 8 * i, 8);
        __builtin_nontemporal_store(w, d + i);
    }
}

// Per-system scalar state lives in two slots.  Kernels come in two roles
// that alternate in every launch sequence: role 0 reads slot 1 and writes
// slot 0, role 1 reads slot 0 and writes slot 1.  A kernel never writes
// what it reads, so no block can see a half-updated state; the block that
// owns a system's first tile is the one that writes the system's new slot
// (every block derives the same values from the same partial sums).
// Nothing crosses between blocks inside a kernel: no atomics, no fences.

}  // namespace eafp
