// walk_check.cpp -- host-side check of the kernels' work decomposition and
// addressing, for every launch shape the library can choose.
//
// For each batch of systems, cells-per-thread and grid size it replays, on
// the CPU, what every block of an iteration kernel does: the TileWalk over
// the tile table, the runs, the cell slots of each tile (with the clamped
// index of slots past a system's last cell), the six neighbour offsets, the
// partial-sum slots that store_run writes and the tile ranges reduce_tiles
// reads.  It checks that
//   - every address a kernel forms lies inside its own system's cells,
//   - every cell is computed by exactly one thread of one block,
//   - decompose() recovers (i, j, k) exactly and neighbours() gives the
//     stencil (theta wrap, r/z walls flagged),
//   - every tile's partial slot is written exactly once, by its system,
//   - a run never crosses a system or a chunk.
// For batches whose every nz is a multiple of 4 it also replays the quad
// kernels (a thread takes 4 consecutive cells of a z-row): every 16-byte
// vector access must start 16-byte aligned and lie inside its system, and
// every cell must be covered exactly once.
// It replays the column-marching stencils (quad and generic) the same way (column table, quad
// and halo lanes, planes i0-1 .. i1 plus the prefetch of i+1 and i+2, the
// LDS slab indices, the partial slots) and checks that every global access
// stays inside its system, every LDS index inside its array, every cell is
// computed exactly once and every partial slot written exactly once.
// It replays the z-line preconditioned kernels too (check_lines).
// Prints "ok <cases>" and exits 0, or the first violation and exits 1.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../hornofusion-full3d_amd/csrc/eafp_common.hpp"

using namespace eafp;

static long g_cases = 0;

#define CHECK(cond, ...)                                                   \
    do {                                                                   \
        if (!(cond)) {                                                     \
            std::printf("FAIL %s:%d: %s: ", __FILE__, __LINE__, #cond);    \
            std::printf(__VA_ARGS__);                                      \
            std::printf("\n");                                             \
            std::exit(1);                                                  \
        }                                                                  \
    } while (0)

static void check_shape(std::vector<SysGeom> g, int cpt, int G) {
    const long ncells = layout_systems(g);
    const int TILE = BLOCK * cpt;
    std::vector<TileDesc> tiles = make_tile_table(g, TILE);
    const int ntiles = (int)tiles.size();
    std::vector<unsigned char> hit(ncells, 0), part(ntiles, 0);
    for (int b = 0; b < G; ++b) {
        const TileWalk w(ntiles, G, b);
        CHECK(w.end <= ntiles && w.stride == G / NXCD, "b=%d", b);
        int prev = -1;
        for (int t = w.first; t < w.end;) {
            CHECK(t > prev, "walk not increasing b=%d t=%d", b, t);
            prev = t;
            const TileDesc T = tiles[t];
            const SysGeom& S = g[T.sys];
            CHECK(T.tile_begin == S.tile_begin && T.tile_end == S.tile_end && t >= T.tile_begin && t < T.tile_end,
                  "tile %d", t);
            CHECK(T.tile_begin >= 0 && T.tile_end <= ntiles, "reduce range of tile %d", t);
            const int nrun = w.run(t, T.tile_end);
            CHECK(nrun >= 1, "empty run b=%d t=%d", b, t);
            for (int u = 0; u < nrun; ++u) {
                const int tile = t + u * w.stride;
                CHECK(tile < w.end && tile < T.tile_end && tiles[tile].sys == T.sys, "run leaves system b=%d t=%d",
                      b, tile);
                CHECK(part[tile] == 0, "partial slot %d written twice", tile);
                part[tile] = 1;
                for (int tid = 0; tid < BLOCK; ++tid) {
                    const int c0 = (tile - T.tile_begin) * TILE + tid;
                    for (int m = 0; m < cpt; ++m) {
                        const int c = c0 + m * BLOCK;
                        const bool ok = c < T.n;
                        const int cc = clamp_cell(c, T.n);
                        CHECK(cc >= 0 && cc < T.n, "slot c=%d n=%d", c, T.n);
                        const CellIdx id = decompose(cc, T);
                        CHECK(id.i >= 0 && id.i < T.nr && id.j >= 0 && id.j < T.nth && id.k >= 0 && id.k < T.nz,
                              "decompose %d -> %d %d %d", cc, id.i, id.j, id.k);
                        CHECK(((long)id.i * T.nth + id.j) * T.nz + id.k == cc, "decompose inverse %d", cc);
                        const long o = T.off + cc;
                        const Nb nb = neighbours(id, T, o);
                        const long lo = T.off, hi = T.off + T.n;
                        const long idx[6] = {nb.W, nb.E, nb.S, nb.N, nb.B, nb.T};
                        for (long v : idx) CHECK(v >= lo && v < hi, "neighbour %ld outside [%ld,%ld)", v, lo, hi);
                        const long sr = (long)T.nth * T.nz;
                        CHECK(nb.hW == (id.i > 0) && nb.W == (nb.hW ? o - sr : o), "W of %d", cc);
                        CHECK(nb.hE == (id.i < T.nr - 1) && nb.E == (nb.hE ? o + sr : o), "E of %d", cc);
                        CHECK(nb.hB == (id.k > 0) && nb.B == (nb.hB ? o - 1 : o), "B of %d", cc);
                        CHECK(nb.hT == (id.k < T.nz - 1) && nb.T == (nb.hT ? o + 1 : o), "T of %d", cc);
                        const int js = (id.j + T.nth - 1) % T.nth, jn = (id.j + 1) % T.nth;
                        CHECK(nb.S == T.off + ((long)id.i * T.nth + js) * T.nz + id.k, "S of %d", cc);
                        CHECK(nb.N == T.off + ((long)id.i * T.nth + jn) * T.nz + id.k, "N of %d", cc);
                        if (ok) {
                            CHECK(hit[o] == 0, "cell %ld computed twice", o);
                            hit[o] = 1;
                        }
                    }
                }
            }
            t += nrun * w.stride;
            CHECK(t >= w.end || tiles[t].sys != T.sys, "next run continues system b=%d t=%d", b, t);
        }
    }
    for (const SysGeom& S : g)
        for (long c = 0; c < S.n; ++c) CHECK(hit[S.off + c] == 1, "cell %ld of system at %ld missed", c, S.off);
    for (int t = 0; t < ntiles; ++t) CHECK(part[t] == 1, "partial slot %d never written", t);
    ++g_cases;
}

static void check_quads(std::vector<SysGeom> g, int cpt, int G) {
    const long ncells = layout_systems(g);
    const int TILE = BLOCK * cpt;
    const int Q = cpt / 4;
    std::vector<TileDesc> tiles = make_tile_table(g, TILE);
    const int ntiles = (int)tiles.size();
    std::vector<unsigned char> hit(ncells, 0);
    for (int b = 0; b < G; ++b) {
        const TileWalk w(ntiles, G, b);
        for (int t = w.first; t < w.end;) {
            const TileDesc T = tiles[t];
            CHECK(T.nz % 4 == 0 && T.n % 4 == 0 && T.off % 4 == 0, "quad geometry");
            const int nrun = w.run(t, T.tile_end);
            for (int u = 0; u < nrun; ++u) {
                const int tile = t + u * w.stride;
                for (int tid = 0; tid < BLOCK; ++tid) {
                    const int c0 = (tile - T.tile_begin) * TILE + 4 * tid;
                    for (int m = 0; m < Q; ++m) {
                        const int c = c0 + 4 * m * BLOCK;
                        const bool ok = c < T.n;
                        const int cc = ok ? c : T.n - 4;
                        const CellIdx id = decompose(cc, T);
                        CHECK(id.k % 4 == 0 && id.k + 3 < T.nz, "quad crosses a row at %d", cc);
                        const long o = T.off + cc;
                        const Nb4 nb = neighbours4(id, T, o);
                        const long lo = T.off, hi = T.off + T.n;
                        const long vec[5] = {o, nb.W, nb.E, nb.S, nb.N};
                        for (long v : vec)
                            CHECK(v % 4 == 0 && v >= lo && v + 4 <= hi, "vector %ld outside [%ld,%ld)", v, lo, hi);
                        CHECK(nb.B >= lo && nb.B < hi && nb.T >= lo && nb.T < hi, "B/T of quad %d", cc);
                        const long sr = (long)T.nth * T.nz;
                        CHECK(nb.W == (id.i > 0 ? o - sr : o) && nb.E == (id.i < T.nr - 1 ? o + sr : o), "W/E");
                        CHECK(nb.B == (id.k > 0 ? o - 1 : o) && nb.T == (id.k + 4 < T.nz ? o + 4 : o), "B/T");
                        if (ok)
                            for (int q = 0; q < 4; ++q) {
                                CHECK(hit[o + q] == 0, "cell %ld computed twice", o + q);
                                hit[o + q] = 1;
                            }
                    }
                }
            }
            t += nrun * w.stride;
        }
    }
    for (const SysGeom& S : g)
        for (long c = 0; c < S.n; ++c) CHECK(hit[S.off + c] == 1, "quad cell %ld of system at %ld missed", c, S.off);
    ++g_cases;
}

static void check_columns(std::vector<SysGeom> g, int cpt, int splits) {
    const long ncells = layout_systems(g);
    std::vector<TileDesc> tiles = make_tile_table(g, BLOCK * cpt);
    std::vector<ColDesc> cols;
    if (!make_col_table(g, splits, cols)) return;      // not eligible: the library refuses it too
    CHECK(cols.size() % NXCD == 0, "column table not padded");
    const int ntiles = (int)tiles.size();
    std::vector<unsigned char> hit(ncells, 0), part(ntiles, 0);
    const int SP = COL_CELLS + 2 * COL_MAX_NZ, SW = COL_CELLS + COL_MAX_NZ;
    const int G = (int)cols.size();
    std::vector<unsigned char> seen(G, 0);
    for (int bidx = 0; bidx < G; ++bidx) {
        const int lb = (bidx % NXCD) * (G / NXCD) + bidx / NXCD;
        CHECK(lb >= 0 && lb < G && !seen[lb], "column mapping");
        seen[lb] = 1;
        const ColDesc C = cols[lb];
        if (C.tile < 0) continue;
        const TileDesc T = tiles[C.tile];
        const long lo = T.off, hi = T.off + T.n, sr = (long)T.nth * T.nz;
        const int Qr = T.nz / 4;
        CHECK(C.nj >= 1 && C.nj * T.nz <= COL_CELLS && C.j0 + C.nj <= T.nth && C.i0 < C.i1 && C.i1 <= T.nr,
              "column shape");
        // the geometry the column descriptor carries is its system's
        CHECK(C.sys == T.sys && C.off == T.off && C.nr == T.nr && C.nth == T.nth && C.nz == T.nz &&
                  C.tile == T.tile_begin && C.tile_end == T.tile_end && C.lt_begin == T.lt_begin &&
                  C.lt_end == T.lt_end,
              "column descriptor geometry");
        CHECK(C.dir == 1 || C.dir == -1, "march direction");
        auto base = [&](int i) { return T.off + (long)(i < 0 ? 0 : i < T.nr ? i : T.nr - 1) * sr; };
        const int ib = C.dir > 0 ? C.i0 : C.i1 - 1, np = C.i1 - C.i0;
        for (int tid = 0; tid < BLOCK; ++tid) {
            const int row = tid / Qr, kq = tid - row * Qr;
            const bool act = row < C.nj;
            const long qoff = (long)(C.j0 + (act ? row : 0)) * T.nz + 4 * (act ? kq : 0);
            const bool hal = tid < 2 * Qr;
            const int hrow = tid < Qr ? -1 : C.nj;
            const int hk = hal ? 4 * (tid < Qr ? tid : tid - Qr) : 0;
            const long hoff = (long)(((C.j0 + hrow) % T.nth + T.nth) % T.nth) * T.nz + hk;
            auto vec = [&](long a) { CHECK(a % 4 == 0 && a >= lo && a + 4 <= hi, "column vector %ld", a); };
            // one step's loads (ColSet): z, p_old of i+dir; the face towards i+dir, w1, w2, e of i; halos of i
            auto issue = [&](int i) {
                vec(base(i + C.dir) + qoff);
                vec(base(C.dir > 0 ? i : i - 1) + qoff);
                vec(base(i) + qoff);
                vec(base(i) + hoff);
            };
            // prologue: planes ib-dir and ib, the face between them, the first step's set
            vec(base(ib - C.dir) + qoff);
            vec(base(ib) + qoff);
            vec(base(C.dir > 0 ? ib - 1 : ib) + qoff);
            issue(ib);
            for (int t = 0; t < np; ++t) {
                const int i = ib + C.dir * t;
                CHECK(i >= C.i0 && i < C.i1, "march plane");
                issue(t + 1 < np ? i + C.dir : i);
                if (!act) {
                    // stores of lanes past the slab go to the sink past the arenas
                    CHECK(ncells % 4 == 0, "sink alignment");
                }
                if (act) {
                    for (int q = 0; q < 4; ++q) {
                        const int sidx = (row + 1) * T.nz + 4 * kq + q;
                        CHECK(sidx >= 0 && sidx + T.nz < SP && sidx - T.nz >= 0, "Sp index");
                        CHECK(row * T.nz + 4 * kq + q < SW && (row + 1) * T.nz + 4 * kq + q < SW, "Sw1 index");
                    }
                    const long o = base(i) + qoff;
                    vec(o);
                    for (int q = 0; q < 4; ++q) {
                        CHECK(hit[o + q] == 0, "column cell %ld computed twice", o + q);
                        hit[o + q] = 1;
                    }
                }
                if (hal) CHECK((hrow + 1) * T.nz + hk + 3 < SP, "halo Sp index");
            }
        }
        // one p.q partial per segment, at the system's first tile + seg
        CHECK(C.seg >= 0 && C.seg < C.nseg && T.tile_begin + C.seg < T.tile_end, "segment slot");
        const long ts = T.tile_begin + C.seg;
        CHECK(part[ts] == 0, "column partial slot %ld written twice", ts);
        part[ts] = 1;
    }
    for (const SysGeom& S : g)
        for (long c = 0; c < S.n; ++c) CHECK(hit[S.off + c] == 1, "column cell %ld of system at %ld missed", c, S.off);
    // the update kernels read slots tile_begin .. tile_begin + pa_n - 1: exactly the written ones
    set_partial_counts(g, tiles, splits);
    std::vector<unsigned char> read(ntiles, 0);
    for (const TileDesc& T : tiles) {
        CHECK(T.pa_n == col_segments(g[T.sys], splits), "pa_n of system %d", T.sys);
        for (int u = 0; u < T.pa_n; ++u) read[T.tile_begin + u] = 1;
    }
    for (int t = 0; t < ntiles; ++t) CHECK(part[t] == read[t], "partial slot %d written %d, read %d", t, part[t], read[t]);
    ++g_cases;
}

// The generic column kernel: slab cells tid + m*BLOCK, scalar accesses.
static void check_columns_generic(std::vector<SysGeom> g, int cpt, int splits) {
    const long ncells = layout_systems(g);
    std::vector<TileDesc> tiles = make_tile_table(g, BLOCK * cpt);
    std::vector<ColDesc> cols;
    if (!make_col_table(g, splits, cols)) return;
    std::vector<unsigned char> hit(ncells, 0);
    const int SP = COL_CELLS + 2 * COL_MAX_NZ, SW = COL_CELLS + COL_MAX_NZ;
    const int M = COL_CELLS / BLOCK, MH = 2 * COL_MAX_NZ / BLOCK;
    for (const ColDesc& C : cols) {
        if (C.tile < 0) continue;
        const TileDesc T = tiles[C.tile];
        const long lo = T.off, hi = T.off + T.n, sr = (long)T.nth * T.nz;
        const int nz = T.nz;
        auto base = [&](int i) { return T.off + (long)(i < 0 ? 0 : i < T.nr ? i : T.nr - 1) * sr; };
        const int ib = C.dir > 0 ? C.i0 : C.i1 - 1, np = C.i1 - C.i0;
        CHECK(C.dir == 1 || C.dir == -1, "generic march direction");
        auto in = [&](long a) { CHECK(a >= lo && a < hi, "generic column access %ld outside [%ld,%ld)", a, lo, hi); };
        for (int tid = 0; tid < BLOCK; ++tid) {
            for (int m = 0; m < M; ++m) {
                const int c = tid + m * BLOCK;
                const bool act = c < C.nj * nz;
                const int cc = act ? c : 0;
                const int row = fdiv(cc, nz, T.inv_nz), k = cc - row * nz;
                CHECK(row >= 0 && row < C.nj && k >= 0 && k < nz, "generic slab cell %d", cc);
                const long coff = (long)(C.j0 + row) * nz + k;
                in(base(ib - 1) + coff);
                in(base(ib) + coff);
                in(base(ib + 1) + coff);
                for (int t = 0; t < np; ++t) {
                    const int i = ib + C.dir * t;
                    CHECK(i >= C.i0 && i < C.i1, "generic march plane");
                    if (t + 1 < np) {
                        const int nx = i + C.dir;
                        in(base(i + 2 * C.dir) + coff);
                        in(base(nx) + coff);
                        in(base(C.dir > 0 ? nx : nx - 1) + coff);
                    }
                    if (!act) continue;
                    const int sl = (row + 1) * nz + k;
                    CHECK(sl - nz >= 0 && sl + nz < SP && sl - 1 >= 0 && sl + 1 < SP, "generic Sp index");
                    CHECK(sl < SW && sl - nz >= 0, "generic Sw1 index");
                    CHECK(row * nz + k < COL_CELLS && (k == 0 || row * nz + k - 1 >= 0), "generic Sw2 index");
                    const long o = base(i) + coff;
                    in(o);
                    CHECK(hit[o] == 0, "generic column cell %ld computed twice", o);
                    hit[o] = 1;
                }
            }
            for (int m = 0; m < MH; ++m) {
                const int h = tid + m * BLOCK;
                if (h >= 2 * nz) continue;
                const int hrow = h < nz ? -1 : C.nj;
                const int hk = h < nz ? h : h - nz;
                const int hslot = (hrow + 1) * nz + hk;
                CHECK(hslot >= 0 && hslot < SP && (hslot >= nz || hslot < SW), "generic halo slot");
                const long hoff = (long)(((C.j0 + hrow) % T.nth + T.nth) % T.nth) * nz + hk;
                for (int i = C.i0 - 1; i <= C.i1; ++i) in(base(i) + hoff);
            }
        }
    }
    for (const SysGeom& S : g)
        for (long c = 0; c < S.n; ++c) CHECK(hit[S.off + c] == 1, "generic column cell %ld missed", c);
    ++g_cases;
}

// The z-line preconditioned kernels (k_linefac, k_mupdate_line,
// k_mreplace_line): one line tile per block; thread t of tile L is lane
// j = t%nq of line li = t/nq (nq = ceil(nz/4)) and holds the line's cells
// 4j .. 4j+nv-1 (LineCells: the quad form, every nz a multiple of 4, reads
// off + line0*nz + 4t .. +3 with 16-byte accesses; the generic form reads
// each cell, slots past nv at the lane's first cell); threads past the
// tile's lines take the tile's first cell.  LDS maps fw/bw/lf indexed by t,
// folds over li*nq + m inside the thread's own line; k_linefac's thread li
// walks line line0+li.  Checks every access stays inside its system, quads
// are 16-byte aligned, every cell is covered exactly once, every partial
// slot written once, and the ranges the stencils reduce.
static void check_lines(std::vector<SysGeom> g, int cpt, bool quad) {
    layout_systems(g);
    const int nlt = line_ranges(g);
    const std::vector<TileDesc> tiles = make_tile_table(g, BLOCK * cpt);
    const std::vector<LineTile> lt = make_line_table(g);
    CHECK((int)lt.size() == nlt, "line tiles %zu != %d", lt.size(), nlt);
    long total = 0;
    for (const SysGeom& G : g) total = G.off + G.n > total ? G.off + G.n : total;
    std::vector<int> cover(total, 0), fac_cover(total, 0), slot(nlt, 0);
    for (const TileDesc& T : tiles) {
        CHECK(T.lt_begin == g[T.sys].lt_begin && T.lt_end == g[T.sys].lt_end, "tile range of system %d", T.sys);
    }
    for (int b = 0; b < nlt; ++b) {
        const LineTile L = lt[b];
        const SysGeom& G = g[L.sys];
        CHECK(b >= G.lt_begin && b < G.lt_end, "tile %d outside its system's range", b);
        CHECK(L.tile_begin == G.tile_begin && L.tile_end == G.tile_end, "tile range %d", b);
        CHECK(L.pa_n >= 1 && L.pa_n <= L.tile_end - L.tile_begin && L.pa_n == G.pa_n, "partial count of tile %d", b);
        CHECK(L.nlines >= 1 && L.nlines <= line_lpb(L.nz), "nlines %d", L.nlines);
        CHECK(L.line0 + L.nlines <= G.nr * G.nth, "lines past the system");
        const int nq = line_lanes(L.nz);
        CHECK(nq >= 1 && nq <= BLOCK, "nq %d", nq);
        slot[b] += 1;   // store_run<2>(partB, blockIdx.x, 1, 1, .)
        for (int tid = 0; tid < BLOCK; ++tid) {
            const int li = tid / nq, j = tid - li * nq;
            const bool inl = li < L.nlines;
            long o;
            int nv;
            if (quad) {
                o = L.off + (long)L.line0 * L.nz + 4 * (long)(inl ? tid : 0);
                nv = 4;
                CHECK(o % 4 == 0, "quad alignment");
                CHECK(o >= G.off && o + 4 <= G.off + G.n, "quad outside system (tile %d tid %d)", b, tid);
            } else {
                // the update stages the block's run of cells through LDS (buffers of 4*BLOCK)
                const long run = L.off + (long)L.line0 * L.nz;
                const int n = L.nlines * L.nz;
                CHECK(n <= 4 * BLOCK, "run of %d cells exceeds the LDS buffer", n);
                CHECK(run >= G.off && run + n <= G.off + G.n, "run outside system (tile %d)", b);
                const int c0 = inl ? li * L.nz + 4 * j : 0;
                CHECK(c0 >= 0 && c0 < n, "LDS index %d of %d", c0, n);
                // 16-byte chunks aligned to the arena: [run & ~3, run + n) rounded up to 4
                const int sh = (int)(run & 3), nch = (sh + n + 3) / 4;
                const long gstart = run - sh;
                const long slot_end = G.off + (G.n + 31) / 32 * 32;
                CHECK(G.off % 4 == 0 && gstart >= G.off && gstart + 4L * nch <= slot_end,
                      "staged chunks outside the system's slot (tile %d)", b);
                CHECK(4 * nch <= 4 * BLOCK + 4, "staged chunks exceed the LDS buffer");
                CHECK(sh + c0 + 3 < 4 * BLOCK + 4, "LDS read index");
                o = L.off + (long)(L.line0 + (inl ? li : 0)) * L.nz + 4 * (inl ? j : 0);
                CHECK(o == run + c0, "direct and staged cell differ");
                nv = inl ? (L.nz - 4 * j < 4 ? L.nz - 4 * j : 4) : 1;
                CHECK(nv >= 1, "empty lane");
                for (int k = 0; k < 4; ++k) {
                    const long a = o + (k < nv ? k : 0);
                    CHECK(a >= G.off && a < G.off + G.n, "cell outside system (tile %d tid %d)", b, tid);
                }
                if (inl && j + 1 < nq) CHECK(nv == 4, "partial lane before the line's last");
                if (inl) CHECK(c0 + nv <= n, "LDS cells past the run");
            }
            if (inl) {
                const long c = o - G.off;
                CHECK(c / L.nz == L.line0 + li && c % L.nz == 4 * j, "thread's cells not on its line");
                CHECK(4 * j + nv <= L.nz, "cells past the line");
                for (int k = 0; k < nv; ++k) cover[o + k] += 1;
                for (int m = 0; m < j; ++m) CHECK(li * nq + m < BLOCK, "fw index");
                for (int m = nq - 1; m > j; --m) CHECK(li * nq + m < BLOCK, "bw index");
                if (j + 1 < nq) CHECK(tid + 1 < BLOCK, "lf index");
            }
            if (tid < L.nlines) {   // k_linefac
                const long o0 = L.off + (long)(L.line0 + tid) * L.nz;
                CHECK(o0 >= G.off && o0 + L.nz <= G.off + G.n, "factor line outside system");
                for (int k = 0; k < L.nz; ++k) fac_cover[o0 + k] += 1;
            }
        }
    }
    for (const SysGeom& G : g)
        for (long c = G.off; c < G.off + G.n; ++c) {
            CHECK(cover[c] == 1, "cell %ld covered %d times", c, cover[c]);
            CHECK(fac_cover[c] == 1, "cell %ld factored %d times", c, fac_cover[c]);
        }
    for (int b = 0; b < nlt; ++b) CHECK(slot[b] == 1, "partial slot %d written %d times", b, slot[b]);
    ++g_cases;
}

static SysGeom sys(int nr, int nth, int nz) {
    SysGeom s{};
    s.nr = nr;
    s.nth = nth;
    s.nz = nz;
    return s;
}

int main() {
    std::vector<std::vector<SysGeom>> batches = {
        {sys(60, 120, 84)},                                           // the production p' system
        {sys(1, 1, 1)},                                               // one cell
        {sys(3, 5, 7)},                                               // smaller than one tile
        {sys(20, 36, 28), sys(1, 1, 1), sys(24, 48, 30), sys(7, 3, 2), sys(9, 64, 5)},   // ragged batch
        {sys(60, 120, 84), sys(17, 40, 33), sys(60, 120, 84)},        // production systems with one between
        {sys(2, 1, 300), sys(1, 257, 1), sys(5, 2, 1)},               // degenerate axes
        {sys(120, 240, 168)},                                         // the large single system
        {sys(60, 120, 84), sys(3, 1, 4), sys(1, 7, 8), sys(20, 36, 28)},   // nz multiples of 4, ragged
        {sys(26, 54, 29), sys(18, 36, 50), sys(35, 66, 24), sys(9, 12, 256), sys(4, 5, 1)},   // any nz
        {sys(3, 2, 1024), sys(5, 7, 4), sys(2, 3, 1020), sys(60, 120, 84)},   // z-lines of 1 .. 256 lanes
        {sys(3, 2, 1023), sys(4, 9, 1), sys(2, 5, 2), sys(3, 3, 3), sys(7, 11, 1021)},   // partial last lanes
        {sys(12, 24, 16), sys(20, 36, 28), sys(16, 30, 24), sys(24, 48, 32)},   // the assembled p' meshes
    };
    const int cpts[] = {1, 2, 4, 8};
    for (auto& bt : batches) {
        for (int cpt : cpts) {
            std::vector<SysGeom> g = bt;
            layout_systems(g);
            const int ntiles = (int)make_tile_table(g, BLOCK * cpt).size();
            // every grid rule of the library: one block per tile (rounded to
            // the XCD count), the per-CU candidates of the autotuner on 256
            // CUs, and small grids
            const int rounded = (ntiles + NXCD - 1) / NXCD * NXCD;
            const int grids[] = {8, 16, 24, 136, rounded, rounded + 8, 4 * 256, 8 * 256, 16 * 256};
            bool quad = true;
            for (const SysGeom& s : bt) quad = quad && s.nz % 4 == 0;
            for (int G : grids) {
                if (G > 16 * 256 && G != rounded) continue;
                check_shape(bt, cpt, G);
                if (quad && cpt >= 4) check_quads(bt, cpt, G);
            }
            {
                std::vector<SysGeom> gl = bt;
                layout_systems(gl);
                if (line_eligible(gl)) {
                    if (quad) check_lines(bt, cpt, true);
                    check_lines(bt, cpt, false);
                }
            }
            for (int splits : {1, 2, 3, 4, 8}) {
                if (quad) check_columns(bt, cpt, splits);
                check_columns_generic(bt, cpt, splits);
            }
        }
    }
    std::printf("ok %ld\n", g_cases);
    return 0;
}
