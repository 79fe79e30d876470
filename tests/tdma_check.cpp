// tdma_check.cpp -- host replay of the GPU line TDMA's schedule.
//
// Runs the library's own per-line code (csrc/tdma_line.hpp, the functions the
// kernels call) on the CPU in the order the kernels impose: the pivot pass
// line by line as k_tdma_pivots' threads take them, then each sweep's theta
// copy column by column and each pass one anti-diagonal of the line grid at a
// time, the lines of a diagonal in reverse and in rotated order (the GPU
// gives them no order).  It checks that
//   - every line of every pass is solved exactly once per pass;
//   - the pivot pass writes m and c' on every interior cell of every pass
//     exactly once and nowhere else;
//   - phi, halos included, is bit-identical to the sequential restatement
//     oracle_tdma_3d (oracle/tdma.c) after 1, 2 and 3 sweeps;
// and, built with AddressSanitizer on the host, that no access leaves the
// exact-size arrays.  Prints "ok <cases>" and exits 0, or the first failure
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

// An exact-size heap array (so that the sanitizer sees any overrun).
static double* alloc(long n) {
    double* p = static_cast<double*>(std::malloc(sizeof(double) * (size_t)n));
    CHECK(p != nullptr, "malloc");
    return p;
}

static void run_case(int nr, int nth, int nz, int n_sweep, unsigned seed, bool zero_pivots, int order) {
    const long sy = nr + 4, sz = (long)(nr + 4) * (nth + 4), n = sz * (nz + 4);
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> U(0.1, 1.0), N(-1.0, 1.0);
    double* c[TD_NCOEF];
    for (int q = 0; q < TD_NCOEF; ++q) c[q] = alloc(n);
    for (long o = 0; o < n; ++o) {
        double sum = 0.0;
        for (int q = 0; q < 6; ++q) {
            c[q][o] = U(rng);
            sum += c[q][o];
        }
        c[CP_][o] = sum * (1.0 + 0.3 * U(rng));
        c[CSU][o] = N(rng);
    }
    if (zero_pivots)   // isolated rows with aP = 0: the first pivot or a later one is clamped
        for (long o = 0; o < n; o += 7) {
            c[CP_][o] = (o % 2) ? 0.0 : -0.0;
            c[CW][o] = c[CE][o] = c[CS][o] = c[CN][o] = c[CB][o] = c[CT][o] = 0.0;
        }
    double* phi0 = alloc(n);
    for (long o = 0; o < n; ++o) phi0[o] = N(rng);

    // the sequential restatement
    double* want = alloc(n);
    std::memcpy(want, phi0, sizeof(double) * n);
    CHECK(oracle_tdma_3d(nr, nth, nz, c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7], want, n_sweep) == 0, "oracle");

    // the replay
    TdSys s;
    s.nr = nr; s.nth = nth; s.nz = nz; s.pad_ = 0;
    s.sy = sy; s.sz = sz; s.n = n;
    for (int q = 0; q < TD_NCOEF; ++q) s.c[q] = c[q];
    s.phi = alloc(n);
    std::memcpy(s.phi, phi0, sizeof(double) * n);
    // the kernels address the six pivot arrays as one block of 6n: replay with
    // one allocation so the offsets are the kernels' own
    double* pivblock = alloc(6 * n);
    for (long o = 0; o < 6 * n; ++o) pivblock[o] = std::nan("");
    s.piv = pivblock;
    s.dp = alloc(n);

    // pivot pass: thread t of the flat line index, past the end included
    const long lines = (long)nth * nz + (long)nr * nz + (long)nr * nth;
    std::vector<int> seen((size_t)lines, 0);
    for (long t = 0; t < lines + 300; ++t) {
        int p, u, v;
        if (!pivot_line(s, t, p, u, v)) {
            CHECK(t >= lines, "line %ld dropped", t);
            continue;
        }
        CHECK(t < lines, "thread %ld past the last line got a line", t);
        const int nu = p == 0 ? nth : nr, nv = p == 2 ? nth : nz;
        CHECK(u >= 1 && u <= nu && v >= 1 && v <= nv, "pivot line (%d,%d) of pass %d out of range", u, v, p);
        line_pivots(s, p, u, v);
        seen[(size_t)t]++;
    }
    for (long t = 0; t < lines; ++t) CHECK(seen[(size_t)t] == 1, "line %ld solved %d times", t, seen[(size_t)t]);
    for (int p = 0; p < 3; ++p)
        for (int k = -1; k <= nz + 2; ++k)
            for (int j = -1; j <= nth + 2; ++j)
                for (int i = -1; i <= nr + 2; ++i) {
                    const long o = tat(s, i, j, k);
                    const bool interior = i >= 1 && i <= nr && j >= 1 && j <= nth && k >= 1 && k <= nz;
                    for (int h = 0; h < 2; ++h) {
                        const double m = pivblock[(2 * p + h) * n + o];
                        CHECK(interior == !std::isnan(m), "pivot array %d at (%d,%d,%d): interior %d value %g",
                              2 * p + h, i, j, k, (int)interior, m);
                    }
                }

    // sweeps, diagonal by diagonal
    const double* cc[TD_NCOEF];
    for (int q = 0; q < TD_NCOEF; ++q) cc[q] = c[q];
    for (int sweep = 0; sweep < n_sweep; ++sweep) {
        for (int p = 0; p < 3; ++p) {
            // theta copy, columns in reverse order
            for (int t = (nr + 4) * (nz + 4) - 1; t >= 0; --t) theta_copy_column(s, t % (nr + 4) - 1, t / (nr + 4) - 1);
            const int nu = p == 0 ? nth : nr, nv = p == 2 ? nth : nz;
            std::vector<int> count((size_t)nu * nv, 0);
            for (int w = 0; w <= nu + nv - 2; ++w) {
                int v_lo, v_hi;
                diagonal_range(w, nu, nv, v_lo, v_hi);
                CHECK(v_lo <= v_hi, "empty diagonal %d", w);
                const int m = v_hi - v_lo + 1;
                for (int r = 0; r < m; ++r) {
                    const int idx = order == 0 ? m - 1 - r : (r + w) % m;
                    const int v = v_lo + idx, u = w + 2 - v;
                    CHECK(u >= 1 && u <= nu && v >= 1 && v <= nv, "line (%d,%d) of pass %d", u, v, p);
                    count[(size_t)(u - 1) + (size_t)nu * (v - 1)]++;
                    if (p == 0) solve_line<0>(s, cc, u, v);
                    else if (p == 1) solve_line<1>(s, cc, u, v);
                    else solve_line<2>(s, cc, u, v);
                }
            }
            for (size_t q = 0; q < count.size(); ++q) CHECK(count[q] == 1, "pass %d line %zu solved %d times", p, q, count[q]);
        }
    }
    long bad = 0;
    for (long o = 0; o < n; ++o)
        if (std::memcmp(&s.phi[o], &want[o], sizeof(double)) != 0) ++bad;
    CHECK(bad == 0, "%dx%dx%d sweeps %d seed %u: %ld of %ld padded values differ from the oracle", nr, nth, nz,
          n_sweep, seed, bad, n);
    for (int q = 0; q < TD_NCOEF; ++q) std::free(c[q]);
    std::free(phi0);
    std::free(want);
    std::free(s.phi);
    std::free(pivblock);
    std::free(s.dp);
    ++g_cases;
}

int main() {
    const int shapes[][3] = {{5, 7, 6}, {1, 5, 4}, {6, 1, 3}, {4, 6, 1}, {3, 2, 5}, {1, 1, 1}, {2, 1, 1},
                             {1, 1, 7}, {12, 24, 16}, {9, 2, 7}, {7, 12, 9}, {20, 36, 28}, {60, 120, 84}};
    unsigned seed = 1;
    for (const auto& sh : shapes) {
        const bool big = sh[0] * sh[1] * sh[2] > 100000;
        for (int n_sweep = 1; n_sweep <= (big ? 1 : 3); ++n_sweep)
            for (int order = 0; order < 2; ++order) {
                run_case(sh[0], sh[1], sh[2], n_sweep, seed++, false, order);
                if (!big) run_case(sh[0], sh[1], sh[2], n_sweep, seed++, true, order);
            }
    }
    std::printf("ok %ld\n", g_cases);
    return 0;
}
