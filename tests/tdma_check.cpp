// tdma_check.cpp -- host replay of the GPU line TDMA.
//
// Runs the library's own per-line, per-element and per-cell code
// (csrc/tdma_line.hpp, the functions the kernels call) on the CPU in the
// order the kernels impose:
//   - k_tdma_setup: every line of every pass (thread = line index);
//   - per sweep and pass, k_tdma_pass: the snapshot of the wrap lines, then
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

static void run_case(int nr, int nth, int nz, int n_sweep, unsigned seed, bool zero_pivots, int order) {
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
        lines += p.NL;
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
        replay_pass<0>(s, order);
        for (long cc = ncells - 1; cc >= 0; --cc) move_cell(s, 0, 1, cc);
        replay_pass<1>(s, order);
        for (long cc = 0; cc < ncells; ++cc) move_cell(s, 1, 2, cc);
        replay_pass<2>(s, order);
        for (long cc = ncells - 1; cc >= 0; --cc) move_cell(s, 2, sweep + 1 < n_sweep ? 0 : 3, cc);
    }
    if (n_sweep > 0)
        for (long t = (long)(nr + 4) * (nz + 4) - 1; t >= 0; --t) wall_halo_column(s, t);
    long bad = 0;
    for (long o = 0; o < n; ++o)
        if (std::memcmp((const void*)&s.home[o], &want[o], sizeof(double)) != 0) ++bad;
    CHECK(bad == 0, "%dx%dx%d sweeps %d seed %u: %ld of %ld padded values differ from the oracle", nr, nth, nz,
          n_sweep, seed, bad, n);
    for (void* p : owned) std::free(p);
    for (int q = 0; q < TD_NCOEF; ++q) std::free(c[q]);
    std::free(phi0);
    std::free(want);
    std::free((void*)s.home);
    ++g_cases;
}

int main() {
    // the last shapes give diagonals longer than the staging arrays (groups) and lines of one element
    const int shapes[][3] = {{5, 7, 6},   {1, 5, 4},    {6, 1, 3},   {4, 6, 1},    {3, 2, 5},
                             {1, 1, 1},   {2, 1, 1},    {1, 1, 7},   {12, 24, 16}, {9, 2, 7},
                             {7, 12, 9},  {20, 36, 28}, {60, 120, 84}, {300, 6, 280}, {2, 3000, 2},
                             {1, 90, 90}};
    unsigned seed = 1;
    for (const auto& sh : shapes) {
        const bool big = (long)sh[0] * sh[1] * sh[2] > 100000;
        for (int n_sweep = 1; n_sweep <= (big ? 1 : 3); ++n_sweep)
            for (int order = 0; order < (big ? 1 : 2); ++order) {
                run_case(sh[0], sh[1], sh[2], n_sweep, seed++, false, order);
                if (!big) run_case(sh[0], sh[1], sh[2], n_sweep, seed++, true, order);
            }
    }
    std::printf("ok %ld\n", g_cases);
    return 0;
}
