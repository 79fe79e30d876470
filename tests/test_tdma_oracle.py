"""Pins the oracle's restatement of the line TDMA (oracle/tdma.c:
``tdma``, src/mod_solver_3d.f90:22-46, and ``tdma_3d_mpi``, :252-366) with
checks that do not retype its loops:

* the reference's own known-answer test (tests/unit/test_tdma.f90): x = sin(i),
  b = 4, a = c = 1, n = 50, max error below 1e-12;
* the Thomas algorithm against scipy's banded LU on random diagonally
  dominant systems of many lengths;
* the pivot clamp: |pivot| < 1e-30 becomes sign(1e-30, pivot), shown by the
  exact results it forces;
* one and two full sweeps against an independent sparse statement of the same
  schedule: each pass is one block Gauss-Seidel step whose implicit couplings
  are exactly the interior neighbours on the same line or on a line earlier in
  the pass's loop order, every other neighbour (later lines, halos, the theta
  wrap through the halo copy) taken from phi as it stood; theta halos copied
  before each pass and not after the last.  A variant in which the wrap is
  implicit (what a plain periodic sweep would do) is shown to differ;
* many sweeps converge to scipy's direct solve of the periodic system with
  the r/z halos as Dirichlet data.
"""
import numpy as np
import pytest
import scipy.linalg
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle

DIRS = {"aW": (-1, 0, 0), "aE": (1, 0, 0), "aS": (0, -1, 0), "aN": (0, 1, 0), "aB": (0, 0, -1), "aT": (0, 0, 1)}


def test_reference_known_answer():
    n = 50
    xe = np.sin(np.arange(1, n + 1, dtype=np.float64))
    a = np.ones(n)
    b = np.full(n, 4.0)
    c = np.ones(n)
    a[0] = 0.0
    c[-1] = 0.0
    d = np.empty(n)
    d[0] = b[0] * xe[0] + c[0] * xe[1]
    d[-1] = a[-1] * xe[-2] + b[-1] * xe[-1]
    d[1:-1] = a[1:-1] * xe[:-2] + b[1:-1] * xe[1:-1] + c[1:-1] * xe[2:]
    assert np.abs(oracle.tdma(a, b, c, d) - xe).max() < 1e-12


@pytest.mark.parametrize("n", [1, 2, 3, 17, 60, 256])
def test_thomas_matches_banded_lu(n):
    rng = np.random.default_rng(n)
    a = rng.uniform(-1, 1, n)
    c = rng.uniform(-1, 1, n)
    b = np.abs(a) + np.abs(c) + rng.uniform(0.5, 2.0, n)
    b[rng.random(n) < 0.3] *= -1.0            # dominance with either sign
    d = rng.normal(size=n)
    ab = np.zeros((3, n))
    ab[0, 1:] = c[:-1]
    ab[1] = b
    ab[2, :-1] = a[1:]
    x_ref = scipy.linalg.solve_banded((1, 1), ab, d)
    x = oracle.tdma(a, b, c, d)
    assert np.abs(x - x_ref).max() <= 1e-12 * np.abs(x_ref).max()


def test_pivot_clamp():
    # first pivot: |b1| < 1e-30 -> sign(1e-30, b1); a negative tiny b1 gives -1e30
    x = oracle.tdma([0.0], [-1e-40], [0.0], [1.0])
    assert x[0] == 1.0 / -1e-30
    # interior pivot exactly 0 (+0.0) -> +1e-30: dp2 = 1e30, x1 = 1 - 1e30
    x = oracle.tdma([0.0, 1.0], [1.0, 1.0], [1.0, 0.0], [1.0, 2.0])
    assert x[1] == 1.0 / 1e-30
    assert x[0] == 1.0 - 1.0 / 1e-30


# ---------------------------------------------------------------------------
# independent statement of one sweep


def theta_copy(a):
    nth = a.shape[1] - 4
    a[:, 1, :] = a[:, nth + 1, :]
    a[:, 0, :] = a[:, nth, :]
    a[:, nth + 2, :] = a[:, 2, :]
    a[:, nth + 3, :] = a[:, 3, :]


def line_key(pass_, i, j, k, shape):
    nr, nth, _ = shape
    if pass_ == "r":        # lines (j, k), k outer, j inner
        return k * (nth + 1) + j
    if pass_ == "t":        # lines (i, k), k outer, i inner
        return k * (nr + 1) + i
    return j * (nr + 1) + i  # z: lines (i, j), j outer, i inner


def one_pass(coef, phi, shape, pass_, wrap_implicit=False):
    """x_new = M^-1 (Su + explicit terms), M = aP - (implicit couplings)."""
    nr, nth, nz = shape
    n = nr * nth * nz
    idx = lambda i, j, k: (i - 1) + nr * ((j - 1) + nth * (k - 1))
    rows, cols, vals = [], [], []
    rhs = np.zeros(n)
    inline = {"r": ("aW", "aE"), "t": ("aS", "aN"), "z": ("aB", "aT")}[pass_]
    for k in range(1, nz + 1):
        for j in range(1, nth + 1):
            for i in range(1, nr + 1):
                c = idx(i, j, k)
                P = (i + 1, j + 1, k + 1)
                rows.append(c); cols.append(c); vals.append(coef["aP"][P])
                rhs[c] += coef["Su"][P]
                for name, (di, dj, dk) in DIRS.items():
                    ii, jj, kk = i + di, j + dj, k + dk
                    a = coef[name][P]
                    inside = 1 <= ii <= nr and 1 <= jj <= nth and 1 <= kk <= nz
                    if wrap_implicit and not inside and 1 <= ii <= nr and 1 <= kk <= nz:
                        jj = (jj - 1) % nth + 1       # the variant: the wrap seen through the periodic index
                        inside = True
                    implicit = inside and (name in inline or
                                           line_key(pass_, ii, jj, kk, shape) < line_key(pass_, i, j, k, shape))
                    if implicit:
                        rows.append(c); cols.append(idx(ii, jj, kk)); vals.append(-a)
                    else:
                        rhs[c] += a * phi[ii + 1, jj + 1, kk + 1]
    M = sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
    x = spla.spsolve(M.tocsc(), rhs)
    out = phi.copy()
    out[2:nr + 2, 2:nth + 2, 2:nz + 2] = x.reshape((nr, nth, nz), order="F")
    return out


def model_sweeps(coef, phi, shape, n_sweep, wrap_implicit=False):
    phi = np.array(phi, order="F")
    for _ in range(n_sweep):
        theta_copy(phi)
        phi = one_pass(coef, phi, shape, "r", wrap_implicit)
        theta_copy(phi)
        phi = one_pass(coef, phi, shape, "t", wrap_implicit)
        theta_copy(phi)
        phi = one_pass(coef, phi, shape, "z", wrap_implicit)
    return phi


def random_system(shape, seed, dominance=1.0):
    nr, nth, nz = shape
    rng = np.random.default_rng(seed)
    pad = (nr + 4, nth + 4, nz + 4)
    coef = {n: np.asfortranarray(rng.uniform(0.1, 1.0, pad)) for n in DIRS}
    s = sum(coef[n] for n in DIRS)
    coef["aP"] = np.asfortranarray(s * (1.0 + dominance * rng.uniform(0.05, 0.5, pad)))
    coef["Su"] = np.asfortranarray(rng.normal(size=pad))
    phi = np.asfortranarray(rng.normal(size=pad))   # every halo holds data
    return coef, phi


SHAPES = [(5, 7, 6), (1, 5, 4), (6, 1, 3), (4, 6, 1), (3, 2, 5), (2, 3, 2)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("n_sweep", [1, 2])
def test_sweeps_match_block_gauss_seidel_statement(shape, n_sweep):
    coef, phi = random_system(shape, 10 + sum(shape) + n_sweep)
    got = oracle.tdma_3d(coef, phi, shape, n_sweep)
    want = model_sweeps(coef, phi, shape, n_sweep)
    scale = np.abs(want).max()
    assert np.abs(got - want).max() <= 1e-11 * scale      # every padded cell, halos included
    if shape[1] >= 3:
        other = model_sweeps(coef, phi, shape, n_sweep, wrap_implicit=True)
        assert np.abs(got - other).max() > 1e-6 * scale    # the lagged wrap is what is pinned


def test_zero_sweeps_leave_phi():
    coef, phi = random_system((4, 5, 3), 1)
    np.testing.assert_array_equal(oracle.tdma_3d(coef, phi, (4, 5, 3), 0), phi)


def test_converges_to_direct_solve():
    shape = (6, 8, 5)
    nr, nth, nz = shape
    coef, phi = random_system(shape, 3, dominance=2.0)
    got = oracle.tdma_3d(coef, phi, shape, 200)
    n = nr * nth * nz
    idx = lambda i, j, k: (i - 1) + nr * ((j - 1) + nth * (k - 1))
    rows, cols, vals = [], [], []
    rhs = np.zeros(n)
    for k in range(1, nz + 1):
        for j in range(1, nth + 1):
            for i in range(1, nr + 1):
                c = idx(i, j, k)
                P = (i + 1, j + 1, k + 1)
                rows.append(c); cols.append(c); vals.append(coef["aP"][P])
                rhs[c] += coef["Su"][P]
                for name, (di, dj, dk) in DIRS.items():
                    ii, jj, kk = i + di, (j + dj - 1) % nth + 1, k + dk
                    if 1 <= ii <= nr and 1 <= kk <= nz:
                        rows.append(c); cols.append(idx(ii, jj, kk)); vals.append(-coef[name][P])
                    else:                                   # r/z halo: Dirichlet data
                        rhs[c] += coef[name][P] * phi[ii + 1, jj + 1, kk + 1]
    x = spla.spsolve(sp.csr_matrix((vals, (rows, cols)), shape=(n, n)).tocsc(), rhs)
    xi = got[2:nr + 2, 2:nth + 2, 2:nz + 2].reshape(-1, order="F")
    assert np.abs(xi - x).max() <= 1e-10 * np.abs(x).max()
