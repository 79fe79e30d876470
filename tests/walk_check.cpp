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
            for (int G : grids) {
                if (G > 16 * 256 && G != rounded) continue;
                check_shape(bt, cpt, G);
            }
        }
    }
    std::printf("ok %ld\n", g_cases);
    return 0;
}
