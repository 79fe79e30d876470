// tdma_check.cpp -- host replay of the GPU line TDMA.
//
// Runs the library's own per-line, per-element and per-cell code
// (csrc/tdma_line.hpp, the functions the kernels call) on the CPU in the
// order the kernels impose:
//   - k_tdma_pack: every tile; k_tdma_setup: every line of every
//     pass (thread = line index);
//   - per sweep and pass, k_tdma_prep and k_tdma_chain phase by phase
//     (replay_chain, below) where chain_plan admits the pass (with chunks
//     capped at 3 elements, or as planned) and its pivots are in band, else
//     k_tdma_pass: the snapshot of the wrap lines, then
//     each anti-diagonal in groups of lines: all elements staged (threads in
//     reverse order), then the lines' recurrences (in reverse or rotated
//     order: the GPU gives them none);
//   - k_tdma_move between the passes' layouts and into the caller's array,
//     k_tdma_wall_halo;
// with the layout the host code of tdma.hip builds (diagonal starts, halo
// lines, array sizes).  It checks that
//   - every line is set up once, every line solved once per pass, every
//     staging slot written once per group and never past the group;
//   - phi, halos included, is bit-identical to the sequential restatement
//     oracle_tdma_3d (oracle/tdma.c) after 1, 2 and 3 sweeps;
// and, built with AddressSanitizer on the host, that no access leaves its
// exact-size array.  Prints "ok <cases>" and exits 0, or the first failure
// and exits 1.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "../hornofusion-full3d_amd/csrc/tdma_line.hpp"

extern "C" int oracle_tdma_3d(int nr, int nth, int nz, const double* aW, const double* aE, const double* aS,
                              const double* aN, const double* aB, const double* aT, const double* aP,
                              const double* Su, double* phi, int n_sweep);

using namespace eafp::tdma;

static long g_cases = 0, g_chain_passes = 0, g_off_band = 0;

#define CHECK(cond, ...)                                                   \
    do {                                                                   \
        if (!(cond)) {                                                     \
            std::printf("FAIL %s:%d: %s: ", __FILE__, __LINE__, #cond);    \
            std::printf(__VA_ARGS__);                                      \
            std::printf("\n");                                             \
            std::exit(1);                                                  \
        }                                                                  \
    } while (0)

template <typename T>
static T* alloc(long n) {
    T* p = static_cast<T*>(std::malloc(sizeof(T) * (size_t)(n > 0 ? n : 1)));
    CHECK(p != nullptr, "malloc");
    return p;
}

// The layout of tdma.hip's host code.
static void pass_shape(int nr, int nth, int nz, int P, int& nu, int& nv, int& n) {
    if (P == 0) { nu = nth; nv = nz; n = nr; }
    else if (P == 1) { nu = nr; nv = nz; n = nth; }
    else { nu = nr; nv = nth; n = nz; }
}

static int halo_lines(int P, int nu, int nv) { return P == 0 ? 2 * nu + nv : P == 1 ? 2 * nv + 2 * nu : 2 * nv + nu; }

template <int P>
static void replay_pass(const TdSys& s, int order) {
    const TdPass& p = s.ps[P];
    if (P != 1) {
        const long nsnap = (long)p.n * (P == 0 ? p.nv : p.nu);
        for (long t = nsnap - 1; t >= 0; --t) snapshot_value(p, P, t);
    }
    double* sd = alloc<double>(TD_LDS_ELEMS);
    double* sa = alloc<double>(TD_LDS_ELEMS);
    double* sm = alloc<double>(TD_LDS_ELEMS);
    double* sc = alloc<double>(TD_LDS_ELEMS);
    std::vector<int> solved((size_t)p.NL, 0);
    std::vector<int> slot(TD_LDS_ELEMS);
    const int gmax = group_lines(p.n);
    for (int w = 0; w <= p.nu + p.nv - 2; ++w) {
        const DiagCtx dc = diag_ctx(p, w);
        int v_lo, v_hi;
        diagonal_range(w, p.nu, p.nv, v_lo, v_hi);
        const int m = v_hi - v_lo + 1, l0 = p.dstart[w];
        CHECK(m >= 1, "empty diagonal");
        if (w > 0) {
            int plo, phi_;
            diagonal_range(w - 1, p.nu, p.nv, plo, phi_);
            CHECK(p.dstart[w] == p.dstart[w - 1] + (phi_ - plo + 1), "diagonal starts");
        }
        for (int g0 = 0; g0 < m; g0 += gmax) {
            const int gm = m - g0 < gmax ? m - g0 : gmax, cnt = gm * p.n;
            CHECK(cnt <= TD_LDS_ELEMS, "group of %d x %d elements exceeds the staging arrays", gm, p.n);
            std::fill(slot.begin(), slot.end(), 0);
            for (int idx = cnt - 1; idx >= 0; --idx) {
                const int q = idx % gm, e = idx / gm;
                const int v = v_lo + g0 + q, u = w + 2 - v;
                int nb[4], nb_ref[4];
                neighbours_in(p, P, dc, u, v, nb);
                line_neighbours(p, P, u, v, nb_ref);
                for (int d = 0; d < 4; ++d)
                    CHECK(nb[d] == nb_ref[d] && nb[d] >= 0 && nb[d] < p.NLX, "pass %d line (%d,%d) neighbour %d: %d vs %d",
                          P, u, v, d, nb[d], nb_ref[d]);
                stage_elem<P>(p, l0 + g0 + q, e, nb, sd, sa, sm, sc, idx);
                slot[(size_t)idx]++;
            }
            for (int idx = 0; idx < cnt; ++idx) CHECK(slot[(size_t)idx] == 1, "slot %d staged %d times", idx, slot[(size_t)idx]);
            for (int r = 0; r < gm; ++r) {
                const int q = order == 0 ? gm - 1 - r : (r + w) % gm;
                const int l = l0 + g0 + q;
                int u, v;
                line_at(p, l, u, v);
                CHECK(u + v - 2 == w && u >= 1 && u <= p.nu && v >= 1 && v <= p.nv, "line %d is (%d,%d)", l, u, v);
                CHECK(line_index(p, u, v) == l, "line index round trip");
                solve_staged(p, l, sd, sa, sm, sc, q, gm);
                solved[(size_t)l]++;
            }
            for (int idx = cnt - 1; idx >= 0; --idx) store_staged(p, l0 + g0, gm, sd, idx);
        }
    }
    for (int l = 0; l < p.NL; ++l) CHECK(solved[(size_t)l] == 1, "pass %d line %d solved %d times", P, l, solved[(size_t)l]);
    std::free(sd); std::free(sa); std::free(sm); std::free(sc);
}

// k_tdma_prep then k_tdma_chain for pass P, phase by phase: the helpers'
// work of a phase (fill the ring slot and the c' entries of the next forward
// chunk) and the chain lanes' work (a forward chunk or the whole back
// substitution) run in either order (order 0: helpers first), which they
// may since they touch different ring slots and different elements of the
// x slots.  Every ring slot is filled with NaN before the helpers fill it,
// every loaded record starts as NaN and the c' entries start as NaN, so a
// read of a value not written for it reaches phi; the x slots start as the
// kernel starts them, 1.0, which only a wall or wrap neighbour may read
// (any other read of it changes phi).  Checks that
// the helpers fill exactly the ring entries of their chunk, that they
// prepare every forward chunk once, in the phase before it, and that the
// chain solves every line once.
template <int P>
static void replay_chain(const TdSys& s, int order) {
    const TdPass& p = s.ps[P];
    for (long t = (long)p.n * p.NL - 1; t >= 0; --t) prep_elem<P>(p, (int)(t % p.NL), (int)(t / p.NL));
    const int NC = 64 * p.NW, HL = TDC_BLOCK - NC;
    CHECK(p.ML <= NC && p.NW <= 4, "pass %d: %d slots for %d chain lanes", P, p.ML, NC);
    CHECK(chain_lds_doubles(p) <= TDC_LDS, "pass %d: LDS %ld doubles", P, chain_lds_doubles(p));
    CHECK(p.IPL >= 1 && p.IPL <= TDC_IPL && (long)p.C * p.ML <= (long)p.IPL * HL &&
              (long)p.C * p.ML > (long)(p.IPL - 1) * HL,
          "pass %d: chunks of %d x %d for %d helper lanes, %d entries each", P, p.C, p.ML, HL, p.IPL);
    CHECK((long)(p.K - 1) * p.C < p.n && (long)p.K * p.C >= p.n, "pass %d: chunks %d x %d for %d", P, p.K, p.C, p.n);
    const double qnan = std::nan("");
    // the kernel's LDS layout (chain_lds) in one exact-size block, all NaN
    const long nlds = chain_lds_doubles(p);
    double* lds = alloc<double>(nlds);
    for (long o = 0; o < nlds; ++o) lds[o] = qnan;
    const TdLds L = chain_lds(p, lds);
    const long ring = L.ring;
    CHECK(L.CS % 2 == 1 && L.LS % 2 == 1 && L.CS >= p.C && L.LS >= p.n && ring == (long)p.ML * L.CS &&
              L.xb + (long)p.ML * L.LS + TDC_PAD == lds + nlds && (double*)L.dr == lds + TDC_PAD &&
              L.cp + (long)(p.ML + 1) * L.LS == L.xb,
          "pass %d: LDS layout", P);
    // the x slots and the row of ones start as 1.0 (the kernel's first step),
    // by threads in reverse order
    for (int tid = TDC_BLOCK - 1; tid >= 0; --tid) chain_init_slots(p, L, tid, TDC_BLOCK);
    for (long o = -L.LS; o < (long)p.ML * L.LS; ++o) CHECK(L.xb[o] == 1.0, "pass %d: x slot %ld not 1.0", P, o);
    double2* dr[2] = {L.dr, L.dr + ring};
    double2* am[2] = {L.am, L.am + ring};
    std::vector<double> carry((size_t)NC, 0.0);
    const long nph = chain_phases(p), nf = (long)(p.nu + p.nv - 1) * p.K;
    TdTask tc = chain_first(p), f = chain_first_fwd(p);
    long prepared = 0, fwd_lines = 0, bwd_lines = 0;
    for (long X = 0; X < nph; ++X) {
        auto helpers = [&]() {
            if (prepared >= nf || chain_prep_phase(p, f) != X) return;
            CHECK(f.kind == 1, "pass %d phase %ld: helpers' task is not a forward chunk", P, X);
            const int bn = (int)((X + 1) & 1);
            for (long o = 0; o < ring; ++o) {
                dr[bn][o] = make_double2(qnan, qnan);
                am[bn][o] = make_double2(qnan, qnan);
            }
            for (int h = HL - 1; h >= 0; --h)
                for (int i = 0; i < p.IPL; ++i) {
                    TdPre x;
                    for (int q = 0; q < RC_N; ++q) x.r[q] = qnan;
                    chain_load<P>(p, f, h + i * HL, x);
                    chain_prepare<P>(p, L, f, bn, h + i * HL, x);
                }
            // the copy of the diagonal before's chunk (the sink on the first diagonal)
            {
                const bool real = f.w > 0;
                int v_lo = 1, hi = 0;
                if (real) diagonal_range(f.w - 1, p.nu, p.nv, v_lo, hi);
                CHECK(!real || (f.pv_lo == v_lo && f.pm == hi - v_lo + 1 && f.pl0 == f.l0 - f.pm &&
                                f.pmM == td_magic((unsigned)f.pm)),
                      "pass %d: the iterator's diagonal before %d", P, f.w);
                const int mp = real ? f.pm : 1;
                CHECK(!real || (long)f.ce * mp <= (long)p.IPL * HL, "pass %d: copy of %d x %d", P, f.ce, mp);
                for (int h = 0; h < HL; ++h)
                    for (int i = 0; i < p.IPL; ++i)
                        chain_copy_out(p, L, real, f.w - 1, f.pv_lo, mp, real ? f.pmM : 0u, f.pl0, f.e0, f.ce,
                                       h + i * HL);
            }
            for (long tt = 0; tt < L.CS; ++tt)
                for (long j = 0; j < p.ML; ++j) {
                    const long o = j * L.CS + tt;
                    const bool in = tt < f.ce && j < f.m;
                    const bool set = !std::isnan(dr[bn][o].x) && !std::isnan(am[bn][o].y);
                    CHECK(in == set, "pass %d phase %ld: ring entry (%ld,%ld) %s", P, X, tt, j,
                          in ? "not filled" : "filled outside the chunk");
                }
            ++prepared;
            chain_advance_fwd(p, f);
        };
        auto chain = [&]() {
            const int b = (int)(X & 1);
            for (int r = 0; r < NC; ++r) {
                const int j = order == 0 ? NC - 1 - r : (int)((r + X) % NC);
                if (tc.kind == 1) chain_forward(p, L, tc, b, j, carry[(size_t)j]);
                else if (tc.kind == 2) chain_backward(p, L, tc, j);
            }
            if (tc.kind == 1 && tc.k == 0) fwd_lines += tc.m;
            if (tc.kind == 2) bwd_lines += tc.m;
        };
        if (order == 0) {
            helpers();
            chain();
        } else {
            chain();
            helpers();
        }
        chain_advance(p, tc);
    }
    CHECK(prepared == nf, "pass %d: %ld of %ld forward chunks prepared", P, prepared, nf);
    {   // the last diagonal's x, after the last barrier
        TdTask t = chain_first(p);
        for (int w = 0; w < p.nu + p.nv - 2; ++w) chain_next_diagonal(p, t);
        for (int it = p.n * t.m - 1; it >= 0; --it)
            chain_copy_out(p, L, true, t.w, t.v_lo, t.m, t.mM, t.l0, 0, p.n, it);
    }
    CHECK(fwd_lines == p.NL && bwd_lines == p.NL, "pass %d: %ld / %ld of %d lines solved", P, fwd_lines, bwd_lines, p.NL);
    std::free(lds);
}

// k_tdma_move<P, Q>: every tile (work-groups and threads in reverse), the
// LDS buffer NaN-filled before the gather.
static void replay_move(const TdSys& s, int P, int Q) {
    std::vector<double> buf(TD_PT3);
    std::vector<int> order(2 * TD_PT * TD_PT), lidx(2 * TD_PT * TD_PT);
    const int NT = 512, TT = TD_PT * TD_PT;
    PackTile t;
    long tiles = 0;
    while (pack_tile_of(s, tiles, t)) ++tiles;
    for (long b = tiles - 1; b >= 0; --b) {
        CHECK(pack_tile_of(s, b, t), "move: tile %ld", b);
        std::fill(order.begin(), order.end(), -1);
        std::fill(lidx.begin(), lidx.end(), -1);
        for (int tid = NT - 1; tid >= 0; --tid) {
            pack_tile_order(s, P, t, order.data(), lidx.data(), tid, NT);
            if (Q < 3) pack_tile_order(s, Q, t, order.data() + TT, lidx.data() + TT, tid, NT);
        }
        std::fill(buf.begin(), buf.end(), std::nan(""));
        for (int tid = NT - 1; tid >= 0; --tid) move_tile_gather(s, P, t, buf.data(), order.data(), lidx.data(), tid, NT);
        for (int tid = NT - 1; tid >= 0; --tid)
            move_tile_scatter(s, P, Q, t, buf.data(), order.data() + TT, lidx.data() + TT, tid, NT);
    }
}

template <int P>
static void run_pass(const TdSys& s, int order) {
    if (s.ps[P].rec && !s.off_band[P]) {
        replay_chain<P>(s, order);
    } else {
        if (s.ps[P].rec) ++g_off_band;
        replay_pass<P>(s, order);
    }
}

// cmax: 0 replays k_tdma_pass for every pass; otherwise the chain kernel
// wherever chain_plan admits the pass, with chunks of at most cmax elements.
static void run_case(int nr, int nth, int nz, int n_sweep, unsigned seed, bool zero_pivots, int order, int cmax,
                     bool huge = false) {
    const long sy = nr + 4, sz = (long)(nr + 4) * (nth + 4), n = sz * (nz + 4);
    const long ncells = (long)nr * nth * nz;
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> U(0.1, 1.0), N(-1.0, 1.0);
    double* c[TD_NCOEF];
    for (int q = 0; q < TD_NCOEF; ++q) c[q] = alloc<double>(n);
    for (long o = 0; o < n; ++o) {
        double sum = 0.0;
        for (int q = 0; q < 6; ++q) {
            c[q][o] = U(rng);
            sum += c[q][o];
        }
        c[CP_][o] = sum * (1.0 + 0.3 * U(rng));
        c[CSU][o] = N(rng);
    }
    if (zero_pivots)   // isolated rows with aP = +-0: the clamp to sign(1e-30, m)
        for (long o = 0; o < n; o += 7) {
            c[CP_][o] = (o % 2) ? 0.0 : -0.0;
            c[CW][o] = c[CE][o] = c[CS][o] = c[CN][o] = c[CB][o] = c[CT][o] = 0.0;
        }
    if (huge)   // pivots near 1e308: outside the chain kernel's band, so k_tdma_pass takes those passes
        for (long o = 0; o < n; o += 37) c[CP_][o] = 4.0e307 * (1.0 + U(rng));
    double* phi0 = alloc<double>(n);
    for (long o = 0; o < n; ++o) phi0[o] = N(rng);
    double* want = alloc<double>(n);
    std::memcpy(want, phi0, sizeof(double) * n);
    CHECK(oracle_tdma_3d(nr, nth, nz, c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7], want, n_sweep) == 0, "oracle");

    TdSys s;
    s.nr = nr; s.nth = nth; s.nz = nz; s.pad_ = 0;
    s.sy = sy; s.sz = sz; s.n = n; s.ncells = ncells;
    for (int q = 0; q < TD_NCOEF; ++q) s.c[q] = (cgdptr)c[q];
    s.home = (gdptr)alloc<double>(n);
    int off_band[3] = {0, 0, 0};
    s.off_band = (decltype(s.off_band))off_band;
    std::memcpy((void*)s.home, phi0, sizeof(double) * n);
    std::vector<void*> owned;
    long lines = 0;
    for (int P = 0; P < 3; ++P) {
        TdPass& p = s.ps[P];
        pass_shape(nr, nth, nz, P, p.nu, p.nv, p.n);
        p.NL = p.nu * p.nv;
        p.NH = halo_lines(P, p.nu, p.nv);
        p.NLX = p.NL + p.NH;
        CHECK((long)p.n * p.NL == ncells, "line grid covers the cells");
        double* pk = alloc<double>(PK_N * ncells);
        double* ln = alloc<double>(LN_N * (long)p.NL);
        double* ph = alloc<double>((long)p.n * p.NLX);
        int* ds = alloc<int>(p.nu + p.nv - 1);
        for (long o = 0; o < (long)p.n * p.NLX; ++o) ph[o] = std::nan("");   // every read must be of a written value
        int acc = 0;
        for (int w = 0; w <= p.nu + p.nv - 2; ++w) {
            int lo, hi;
            diagonal_range(w, p.nu, p.nv, lo, hi);
            ds[w] = acc;
            acc += hi - lo + 1;
        }
        CHECK(acc == p.NL, "diagonals cover the lines");
        p.pk = (gdptr)pk; p.ln = (gdptr)ln; p.phi = (gdptr)ph; p.dstart = (cgiptr)ds;
        owned.insert(owned.end(), {pk, ln, ph, ds});
        p.rec = p.sink = p.pp = nullptr;
        p.uv = nullptr;
        if (cmax > 0 && chain_plan(p, cmax)) {
            double* rec = alloc<double>(RC_REC * ncells);
            for (long o = 0; o < RC_REC * ncells; ++o) rec[o] = std::nan("");
            double* pp = alloc<double>(2 * ncells);
            for (long o = 0; o < 2 * ncells; ++o) pp[o] = std::nan("");
            p.pp = (gdptr)pp;
            owned.push_back(pp);
            double* sink = alloc<double>(TDC_SINK);
            int* uv = alloc<int>(p.NL);
            for (int l = 0; l < p.NL; ++l) uv[l] = -1;
            p.rec = (gdptr)rec;
            p.sink = (gdptr)sink;
            p.uv = (decltype(p.uv))uv;
            owned.push_back(rec);
            owned.push_back(sink);
            owned.push_back(uv);
            ++g_chain_passes;
        }
        lines += p.NL;
    }
    // k_tdma_pack: every tile (work-groups and threads in reverse), the LDS
    // buffer NaN-filled before each source is staged
    {
        std::vector<double> buf(TD_PT3);
        std::vector<int> order(3 * TD_PT * TD_PT), lidx(3 * TD_PT * TD_PT);
        const int NT = 512;
        PackTile t;
        long tiles = 0;
        while (pack_tile_of(s, tiles, t)) ++tiles;
        CHECK(tiles == (long)((nr + TD_PT - 1) / TD_PT) * ((nth + TD_PT - 1) / TD_PT) * ((nz + TD_PT - 1) / TD_PT),
              "pack: %ld tiles", tiles);
        for (long b = tiles - 1; b >= 0; --b) {
            CHECK(pack_tile_of(s, b, t), "pack: tile %ld", b);
            std::fill(order.begin(), order.end(), -1);
            std::fill(lidx.begin(), lidx.end(), -1);
            for (int P = 0; P < 3; ++P) {
                int nu, nv, ne;
                pack_tile_extents(P, t, nu, nv, ne);
                for (int tid = NT - 1; tid >= 0; --tid)
                    pack_tile_order(s, P, t, order.data() + P * TD_PT * TD_PT, lidx.data() + P * TD_PT * TD_PT, tid, NT);
                std::vector<int> seen((size_t)nu * nv, 0);
                for (int r = 0; r < nu * nv; ++r) {
                    const int o = order[(size_t)P * TD_PT * TD_PT + r], lu = o & 255, lv = o >> 8;
                    CHECK(o >= 0 && lu < nu && lv < nv && ++seen[(size_t)lv * nu + lu] == 1, "pack: order %d of pass %d", r, P);
                    const int u0 = P == 0 ? t.j0 : t.i0, v0 = P == 2 ? t.j0 : t.k0;
                    CHECK(lidx[(size_t)P * TD_PT * TD_PT + r] == line_index(s.ps[P], u0 + lu + 1, v0 + lv + 1),
                          "pack: line index %d of pass %d", r, P);
                    if (r > 0) {   // diagonal by diagonal, v ascending within a diagonal
                        const int p = order[(size_t)P * TD_PT * TD_PT + r - 1], pu = p & 255, pv = p >> 8;
                        CHECK(pu + pv < lu + lv || (pu + pv == lu + lv && pv < lv), "pack: order of pass %d at %d", P, r);
                    }
                }
            }
            for (int src = 0; src < TD_PSRC; ++src) {
                std::fill(buf.begin(), buf.end(), std::nan(""));
                for (int tid = NT - 1; tid >= 0; --tid) pack_tile_load(s, t, src, buf.data(), tid, NT);
                for (int P = 0; P < 3; ++P)
                    for (int tid = NT - 1; tid >= 0; --tid)
                        pack_tile_store(s, P, t, src, buf.data(), order.data() + P * TD_PT * TD_PT,
                                        lidx.data() + P * TD_PT * TD_PT, tid, NT);
            }
        }
    }
    // k_tdma_setup: thread t = line index within the passes, past the end included
    std::vector<int> setup_seen((size_t)lines, 0);
    for (long t0 = lines + 40; t0 >= 0; --t0) {
        long t = t0;
        for (int P = 0; P < 3; ++P) {
            const TdPass& p = s.ps[P];
            if (t < p.NL) {
                int u, v;
                line_at(p, (int)t, u, v);
                CHECK(line_index(p, u, v) == (int)t, "setup line round trip");
                setup_line(s, P, u, v);
                setup_seen[(size_t)t0]++;
                break;
            }
            t -= p.NL;
        }
    }
    for (long t = 0; t < lines; ++t) CHECK(setup_seen[(size_t)t] == 1, "setup line %ld seen %d times", t, setup_seen[(size_t)t]);
    for (int sweep = 0; sweep < n_sweep; ++sweep) {
        run_pass<0>(s, order);
        replay_move(s, 0, 1);
        run_pass<1>(s, order);
        replay_move(s, 1, 2);
        run_pass<2>(s, order);
        replay_move(s, 2, sweep + 1 < n_sweep ? 0 : 3);
    }
    if (n_sweep > 0)
        for (long t = (long)(nr + 4) * (nz + 4) - 1; t >= 0; --t) wall_halo_column(s, t);
    long bad = 0;
    for (long o = 0; o < n; ++o)
        if (std::memcmp((const void*)&s.home[o], &want[o], sizeof(double)) != 0) ++bad;
    CHECK(bad == 0, "%dx%dx%d sweeps %d seed %u cmax %d: %ld of %ld padded values differ from the oracle", nr, nth, nz,
          n_sweep, seed, cmax, bad, n);
    for (void* p : owned) std::free(p);
    for (int q = 0; q < TD_NCOEF; ++q) std::free(c[q]);
    std::free(phi0);
    std::free(want);
    std::free((void*)s.home);
    ++g_cases;
}

// td_divmod against `/` and `%`: every divisor up to 1024, every dividend
// below 2^20 / m (the kernels' dividends stay below 2^18, divisors below
// 2^10).
static void check_divmod() {
    for (int m = 1; m <= 1024; ++m) {
        const unsigned M = td_magic((unsigned)m);
        for (int n = 0; n < (1 << 20) / m; ++n) {
            int q, r;
            td_divmod(n, m, M, q, r);
            CHECK(q == n / m && r == n % m, "td_divmod(%d, %d): %d %d", n, m, q, r);
        }
    }
}

int main() {
    check_divmod();
    // the last shapes give diagonals longer than the staging arrays (groups) and lines of one element
    const int shapes[][3] = {{5, 7, 6},   {1, 5, 4},    {6, 1, 3},   {4, 6, 1},    {3, 2, 5},
                             {1, 1, 1},   {2, 1, 1},    {1, 1, 7},   {12, 24, 16}, {9, 2, 7},
                             {7, 12, 9},  {20, 36, 28}, {60, 120, 84}, {300, 6, 280}, {2, 3000, 2},
                             {1, 90, 90}};
    unsigned seed = 1;
    // each shape with k_tdma_pass (cmax 0) and the chain kernel, with chunks
    // of at most 3, 7 and 13 elements (several chunks, a short last one; one,
    // two and four groups of the forward's reads) and as planned
    const int cmaxes[] = {0, 3, 7, 13, 1 << 20};
    for (const auto& sh : shapes) {
        const bool big = (long)sh[0] * sh[1] * sh[2] > 100000;
        for (int cmax : cmaxes)
            for (int n_sweep = 1; n_sweep <= (big ? 1 : 3); ++n_sweep)
                for (int order = 0; order < (big ? 1 : 2); ++order) {
                    if (big && cmax > 0 && cmax < 1 << 20) continue;
                    run_case(sh[0], sh[1], sh[2], n_sweep, seed++, false, order, cmax);
                    if (!big) run_case(sh[0], sh[1], sh[2], n_sweep, seed++, true, order, cmax);
                }
    }
    const int huge_shapes[][3] = {{12, 24, 16}, {7, 12, 9}};
    for (const auto& sh : huge_shapes) run_case(sh[0], sh[1], sh[2], 2, seed++, false, 0, 1 << 20, true);
    CHECK(g_off_band > 0, "no pass left the band");
    CHECK(g_chain_passes > 100, "only %ld passes ran the chain kernel", g_chain_passes);
    std::printf("ok %ld %ld\n", g_cases, g_chain_passes);
    return 0;
}
