"""Pins the FP64 oracle (oracle/cg_oracle.c) before it is trusted as the
checker of the GPU path.

No golden vector or known-answer test of cg_3d_mpi exists in the reference
(its tests reach the CG only through solve_pressure_correction), and its
Fortran/MPI build cannot run in this image, so each oracle function is
pinned by what the mathematics fixes, chosen so that a plausible mistake
(wrong neighbour stride, sign, theta wrap, dropped term, wrong alpha/beta,
wrong stop test) fails at least one check:

* matvec   -- an independent scipy CSR operator on random x; the closed form
              A*1 = aP - sum(a_nb); theta-roll equivariance (the periodic
              wrap, src/mod_mpi_topology.f90:303-310).
* theta_halo -- the four plane copies, checked directly.
* dot      -- math.fsum (correctly rounded) within a few ulps.
* cg       -- scipy.sparse.linalg.cg with the Jacobi preconditioner (the same
              textbook PCG recurrences and the same recurrence-residual stop
              test) gives the same iteration count and the same iterates in
              the first iterations; scipy spsolve of the gauge-fixed
              (bordered) system; manufactured solutions b = A x*; invariants
              of the reference tests: Su = 0 returns at once with x = 0
              (test_poisson_bath.f90:107-141), identity rows stay 0
              (test_sealed_cell.f90:64-71), breakdown exit, max_iter cap.
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

import oracle
from eafp import systems


@pytest.fixture(scope="module")
def sys1():
    return systems.make_system(1)


def padded_random(ps, seed, halos=True):
    rng = np.random.default_rng(seed)
    x = np.zeros((ps.nr + 4, ps.nth + 4, ps.nz + 4), order="F")
    x[2:-2, 2:-2, 2:-2] = rng.normal(size=ps.shape)
    return x


def test_matvec_matches_independent_csr(sys1):
    ps = sys1
    x = padded_random(ps, 1)
    y = oracle.matvec(ps, x)
    A = oracle.csr(ps)
    y_ref = A @ oracle.interior(x).ravel()
    np.testing.assert_allclose(oracle.interior(y).ravel(), y_ref, rtol=1e-12, atol=1e-12 * np.abs(y_ref).max())
    # y is zero outside the interior (y = 0 over the whole padded array first)
    yy = y.copy()
    yy[2:-2, 2:-2, 2:-2] = 0.0
    assert not yy.any()


def test_matvec_closed_form_on_constant(sys1):
    ps = sys1
    one = np.ones((ps.nr + 4, ps.nth + 4, ps.nz + 4), order="F")
    y = oracle.interior(oracle.matvec(ps, one))
    expect = ps.interior("aP") - sum(ps.interior(n) for n in ("aW", "aE", "aS", "aN", "aB", "aT"))
    np.testing.assert_allclose(y, expect, atol=1e-12 * ps.interior("aP").max())
    # pure Neumann rows -> 0, identity rows -> 1
    assert np.abs(y[ps.active]).max() < 1e-12 * ps.interior("aP").max()
    np.testing.assert_array_equal(y[~ps.active], 1.0)


def test_matvec_theta_roll_equivariance(sys1):
    ps = sys1
    s = 5
    x = padded_random(ps, 2)
    y = oracle.interior(oracle.matvec(ps, x.copy(order="F")))
    rolled = systems.PSystem(ps.nr, ps.nth, ps.nz, {}, None, ps.active, ps.labels, ps.ncomp, {})
    for n in systems.COEF_NAMES:
        a = np.zeros_like(ps.coef[n])
        a[2:-2, 2:-2, 2:-2] = np.roll(ps.interior(n), s, axis=1)
        rolled.coef[n] = a
    xr = np.zeros_like(x)
    xr[2:-2, 2:-2, 2:-2] = np.roll(oracle.interior(x), s, axis=1)
    yr = oracle.interior(oracle.matvec(rolled, xr))
    np.testing.assert_array_equal(yr, np.roll(y, s, axis=1))


def test_theta_halo_planes():
    shape = (3, 5, 4)
    x = np.asfortranarray(np.random.default_rng(3).normal(size=(7, 9, 8)))
    x0 = x.copy(order="F")
    oracle.theta_halo(shape, x)
    nth = shape[1]
    J = lambda j: j + 1
    np.testing.assert_array_equal(x[:, J(0), :], x0[:, J(nth), :])
    np.testing.assert_array_equal(x[:, J(-1), :], x0[:, J(nth - 1), :])
    np.testing.assert_array_equal(x[:, J(nth + 1), :], x0[:, J(1), :])
    np.testing.assert_array_equal(x[:, J(nth + 2), :], x0[:, J(2), :])
    np.testing.assert_array_equal(x[:, 2:-2, :], x0[:, 2:-2, :])


def test_dot_against_fsum(sys1):
    ps = sys1
    a = padded_random(ps, 4)
    b = padded_random(ps, 5)
    a[0, :, :] = 1e300   # halos must not enter
    d = oracle.dot(ps.shape, a, b)
    ref = oracle.fsum_dot(oracle.interior(a), oracle.interior(b))
    assert abs(d - ref) <= 1e-13 * math.sqrt(oracle.fsum_dot(oracle.interior(a), oracle.interior(a)) *
                                             oracle.fsum_dot(oracle.interior(b), oracle.interior(b)))


@pytest.mark.parametrize("config,shape", [(1, None), (3, (18, 36, 24))])
def test_cg_matches_scipy_pcg(config, shape):
    """scipy's PCG (Jacobi M) runs the same recurrences: same count, same early iterates."""
    ps = systems.make_system(config, 0, shape)
    A = oracle.csr(ps)
    b = ps.interior("Su").ravel()
    Minv = sp.diags(1.0 / np.maximum(np.abs(A.diagonal()), 1e-30))
    xs = []
    x_sc, info = sla.cg(A, b, rtol=1e-8, atol=0.0, M=Minv, maxiter=100000, callback=lambda xk: xs.append(xk.copy()))
    assert info == 0
    phi, res, nit, hist = oracle.cg(ps, max_iter=100000, tol=1e-8, history=True)
    assert abs(nit - len(xs)) <= 1, (nit, len(xs))
    # iterates agree early on (same recurrences; only rounding differs)
    for k in (1, 2, 5, 10):
        phik, _, _, _ = oracle.cg(ps, max_iter=k, tol=0.0)
        np.testing.assert_allclose(oracle.interior(phik).ravel(), xs[k - 1], rtol=1e-8,
                                   atol=1e-10 * np.abs(xs[k - 1]).max())
    # the recurrence residual history matches the true residual of scipy's iterates early on
    for k in (1, 3, 8):
        tr = np.linalg.norm(b - A @ xs[k - 1]) / np.linalg.norm(b)
        assert abs(hist[k] - tr) <= 1e-8 * tr


def test_cg_matches_bordered_direct_solve(sys1):
    """Gauge-fixed direct solve: [A e; e^T 0] [x; l] = [b; 0] with e the active mask."""
    ps = sys1
    assert ps.ncomp == 1
    A = oracle.csr(ps)
    n = A.shape[0]
    e = ps.active.ravel().astype(np.float64)
    K = sp.bmat([[A, sp.csr_matrix(e[:, None])], [sp.csr_matrix(e[None, :]), None]]).tocsc()
    rhs = np.concatenate([ps.interior("Su").ravel(), [0.0]])
    x_direct = sla.spsolve(K, rhs)[:n]
    phi, res, nit, _ = oracle.cg(ps, max_iter=100000, tol=1e-13)
    err = oracle.gauge_error(oracle.interior(phi), x_direct.reshape(ps.shape), ps.labels, ps.ncomp)
    assert err < 1e-8, err


@pytest.mark.parametrize("config", [1, 3])
def test_cg_manufactured_solution(config):
    ps = systems.make_system(config, 1)
    rng = np.random.default_rng(11)
    xstar = oracle.project(rng.normal(size=ps.shape), ps.labels, ps.ncomp)
    A = oracle.csr(ps)
    b = (A @ xstar.ravel()).reshape(ps.shape)
    ps.Su[2:-2, 2:-2, 2:-2] = b
    phi, res, nit, _ = oracle.cg(ps, max_iter=100000, tol=1e-13)
    assert res < 1e-13
    err = oracle.gauge_error(oracle.interior(phi), xstar, ps.labels, ps.ncomp)
    assert err < 1e-7, err


def test_cg_zero_rhs_returns_at_once(sys1):
    ps = systems.make_system(1)
    ps.Su[...] = 0.0
    phi0 = np.zeros((ps.nr + 4, ps.nth + 4, ps.nz + 4), order="F")
    phi, res, nit, _ = oracle.cg(ps, phi0)
    assert res == 0.0 and nit == 0
    assert not phi.any()


def test_cg_identity_rows_stay_zero(sys1):
    phi, res, nit, _ = oracle.cg(sys1)
    assert res < 1e-8
    assert not oracle.interior(phi)[~sys1.active].any()


def test_cg_breakdown_and_cap():
    ps = systems.make_system(1)
    for n in systems.COEF_NAMES:
        ps.coef[n][...] = 0.0
    phi, res, nit, _ = oracle.cg(ps)
    assert nit == 0 and res == 1.0
    ps = systems.make_system(1)
    for k in (1, 4):
        phi, res, nit, hist = oracle.cg(ps, max_iter=k, history=True)
        assert nit == k and len(hist) == k + 1 and res == hist[-1] > 1e-8


def test_cg_final_theta_halo(sys1):
    phi, res, nit, _ = oracle.cg(sys1)
    nth = sys1.nth
    np.testing.assert_array_equal(phi[:, 1, :], phi[:, nth + 1, :])
    np.testing.assert_array_equal(phi[:, nth + 2, :], phi[:, 2, :])


def test_planted_mistake_is_caught():
    """A wrong beta (no Jacobi weighting in r.z) planted in a numpy copy of the
    recurrence is caught by the scipy PCG comparison used above."""
    ps = systems.make_system(1)
    A = oracle.csr(ps)
    b = ps.interior("Su").ravel()
    d = 1.0 / A.diagonal()
    x = np.zeros_like(b)
    r = b.copy()
    z = r * d
    p = z.copy()
    rz = r @ r          # planted mistake: should be r @ z
    for _ in range(5):
        q = A @ p
        a = rz / (p @ q)
        x += a * p
        r -= a * q
        z = r * d
        rzn = r @ r
        p = z + (rzn / rz) * p
        rz = rzn
    xs = []
    sla.cg(A, b, rtol=1e-8, atol=0.0, M=sp.diags(d), maxiter=5, callback=lambda xk: xs.append(xk.copy()))
    assert np.linalg.norm(x - xs[4]) > 1e-3 * np.linalg.norm(xs[4])


@pytest.mark.parametrize("config,shape", [(1, None), (2, (20, 36, 28)), (5, (24, 48, 30))])
def test_openmp_baseline_matches_sequential_oracle(config, shape):
    """The multi-core baseline (cg_omp.c) is the same CG with parallel dot
    reductions: same iteration count within one, same solution, halos the
    periodic copies."""
    ps = systems.make_system(config, 0, shape)
    phi_s, res_s, nit_s, _ = oracle.cg(ps, max_iter=20000, tol=1e-8)
    phi_p, res_p, nit_p = oracle.cg_omp(ps, max_iter=20000, tol=1e-8)
    assert abs(nit_p - nit_s) <= 1
    assert res_p < 1e-8 and res_s < 1e-8
    err = oracle.gauge_error(oracle.interior(phi_p), oracle.interior(phi_s), ps.labels, ps.ncomp)
    assert err < 1e-6
    nth = ps.shape[1]
    np.testing.assert_array_equal(phi_p[:, 1, :], phi_p[:, nth + 1, :])   # j=0 <- j=nth
    np.testing.assert_array_equal(phi_p[:, nth + 3, :], phi_p[:, 3, :])   # j=nth+2 <- j=2
