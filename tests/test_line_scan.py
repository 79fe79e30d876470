"""CPU check of the z-line preconditioner's algebra, as the mixed solver's
kernels apply it (k_linefac, line_solve in hornofusion-full3d_amd/csrc/
mixed.hip), restated in numpy:

  - factors of M = L diag(d) L^T from the tridiagonal part along z
    (diagonal D, off-diagonals -w): l_k = -w_{k-1}/d_{k-1},
    d_k = max(D_k - w_{k-1}^2/d_{k-1}, D_k/16); rows with D = 0 get l = 0
    and 1/d = 0;
  - z = M^-1 r by lanes of 4 cells (nq = ceil(nz/4), the last lane masked):
    each lane composes its cells' forward maps y -> r_k - l_k y into one
    affine map, every lane folds the maps of the lanes before it; then the
    same backwards with z -> u_k - l_{k+1} z.

Pinned against what the mathematics fixes: scipy's banded solve of T when
the floor is inactive (then M = T), a dense solve of the floored M, and M's
symmetry and positive definiteness.  A planted error (the backward step
using l_k instead of l_{k+1}) is shown to fail.
"""
import numpy as np
import pytest
from scipy.linalg import solve_banded


def factors(D, w):
    """l, 1/d as k_linefac computes them (FP64; the kernel then rounds them)."""
    nz = len(D)
    l = np.zeros(nz)
    inv = np.zeros(nz)
    dprev, wprev = 0.0, 0.0
    for k in range(nz):
        lk, dk = 0.0, D[k]
        if k > 0 and dprev > 0.0:
            lk = -wprev / dprev
            dk = D[k] - wprev * wprev / dprev
        dk = max(dk, 0.0625 * D[k])
        wprev = w[k]
        live = D[k] > 0.0
        l[k] = lk if live else 0.0
        inv[k] = 1.0 / dk if live else 0.0
        dprev = dk if live else 0.0
    return l, inv


def line_solve(r, l, inv, backward_bug=False):
    """The lane algorithm of line_solve: 4 cells per lane, maps folded in order."""
    nz = len(r)
    nq = (nz + 3) // 4
    R = np.zeros(4 * nq)
    Lf = np.zeros(4 * nq)
    IV = np.zeros(4 * nq)
    R[:nz], Lf[:nz], IV[:nz] = r, l, inv          # masked slots: r = l = 1/d = 0
    fw = []
    for j in range(nq):
        A, B = 1.0, 0.0
        for k in range(4):
            c = 4 * j + k
            A, B = -Lf[c] * A, R[c] - Lf[c] * B
        fw.append((A, B))
    u = np.zeros(4 * nq)
    for j in range(nq):
        y = 0.0
        for m in range(j):
            y = fw[m][0] * y + fw[m][1]
        for k in range(4):
            c = 4 * j + k
            y = R[c] - Lf[c] * y
            u[c] = y * IV[c]
    lf = [Lf[4 * j] for j in range(nq)]
    bw = []
    for j in range(nq):
        ln = lf[j + 1] if j + 1 < nq else 0.0
        nxt = [Lf[4 * j + k + 1] for k in range(3)] + [ln]
        if backward_bug:
            nxt = [Lf[4 * j + k] for k in range(4)]
        A, B = -nxt[3], u[4 * j + 3]
        for k in (2, 1, 0):
            A, B = -nxt[k] * A, u[4 * j + k] - nxt[k] * B
        bw.append((A, B, nxt))
    z = np.zeros(4 * nq)
    for j in range(nq):
        zin = 0.0
        for m in range(nq - 1, j, -1):
            zin = bw[m][0] * zin + bw[m][1]
        nxt = bw[j][2]
        z[4 * j + 3] = u[4 * j + 3] - nxt[3] * zin
        for k in (2, 1, 0):
            z[4 * j + k] = u[4 * j + k] - nxt[k] * z[4 * j + k + 1]
    return z[:nz]


def tridiag(D, w):
    """T: diagonal D, off-diagonals -w (w[k] couples k and k+1; w[-1] unused)."""
    nz = len(D)
    T = np.diag(D.astype(float))
    for k in range(nz - 1):
        T[k, k + 1] = T[k + 1, k] = -w[k]
    return T


def make_line(nz, rng, margin):
    """A diagonally dominant line: D_k = w_{k-1} + w_k + lateral margin."""
    w = rng.uniform(0.1, 10.0, nz)
    w[-1] = 0.0
    wm = np.concatenate([[0.0], w[:-1]])
    D = wm + w + margin * rng.uniform(0.5, 1.5, nz)
    return D, w


@pytest.mark.parametrize("nz", [1, 2, 3, 4, 5, 7, 8, 21, 28, 29, 84, 85, 168, 1021])
def test_line_solve_equals_banded_solve(nz):
    rng = np.random.default_rng(nz)
    D, w = make_line(nz, rng, margin=0.5)
    l, inv = factors(D, w)
    r = rng.normal(size=nz)
    z = line_solve(r, l, inv)
    ab = np.zeros((3, nz))
    ab[0, 1:] = -w[:-1]
    ab[1] = D
    ab[2, :-1] = -w[:-1]
    zref = solve_banded((1, 1), ab, r)
    assert np.allclose(z, zref, rtol=1e-10, atol=1e-12 * np.abs(zref).max())


def test_floor_keeps_M_positive_definite():
    """An isolated Neumann column (no lateral margin): T is singular, the
    floor makes M = L diag(d) L^T SPD and the lanes apply exactly M^-1."""
    nz = 29
    rng = np.random.default_rng(7)
    D, w = make_line(nz, rng, margin=0.0)
    assert abs(np.linalg.det(tridiag(D, w) / D.max())) < 1e-8    # singular up to rounding
    l, inv = factors(D, w)
    L = np.eye(nz) + np.diag(l[1:], -1)
    M = L @ np.diag(1.0 / inv) @ L.T
    assert np.allclose(M, M.T)
    assert np.linalg.eigvalsh(M).min() > 0
    r = rng.normal(size=nz)
    assert np.allclose(line_solve(r, l, inv), np.linalg.solve(M, r), rtol=1e-9, atol=1e-12)


def test_empty_rows_give_zero():
    """Rows with D = 0 (no links): z = 0 there, as Jacobi's 1/D = 0, and the
    lines on either side decouple."""
    nz = 12
    rng = np.random.default_rng(3)
    D, w = make_line(nz, rng, margin=0.5)
    D[5] = 0.0
    w[4] = w[5] = 0.0
    D[4] = D[4] - 0.0
    l, inv = factors(D, w)
    r = rng.normal(size=nz)
    z = line_solve(r, l, inv)
    assert z[5] == 0.0
    for seg in (slice(0, 5), slice(6, nz)):
        Ds, ws = D[seg], w[seg].copy()
        ws[-1] = 0.0
        zs = np.linalg.solve(tridiag(Ds, ws), r[seg])
        assert np.allclose(z[seg], zs, rtol=1e-10)


def test_planted_backward_index_error_is_caught():
    nz = 21
    rng = np.random.default_rng(11)
    D, w = make_line(nz, rng, margin=0.5)
    l, inv = factors(D, w)
    r = rng.normal(size=nz)
    zref = np.linalg.solve(tridiag(D, w), r)
    assert not np.allclose(line_solve(r, l, inv, backward_bug=True), zref, rtol=1e-6)
