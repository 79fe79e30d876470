"""GPU parity of the FP64 drop-in (eafp_cg_solve / batched FP64 CG) against
the oracle restatement of cg_3d_mpi (src/mod_solver_3d.f90:485-620).

Bar: same stop outcome and iteration count within a few iterations of the
oracle (reduction order and FMA differ, so iteration counts near the
threshold can move by one or two), relative p' error after the zero-mean
gauge fix < 1e-6 at tol 1e-8, true relative residual < 1.5*tol, theta
halos exact periodic copies, r/z halos untouched.  Edge cases: zero RHS
(early return), max_iter = 1, breakdown, NaN input, ragged batches.
"""
import math

import numpy as np
import pytest

import eafp
import oracle
from eafp import systems
from eafp.solver import COEF_NAMES

pytestmark = pytest.mark.gpu


def drop_in(ps, phi, max_iter=100000, tol=1e-8):
    m = eafp.Mesh(ps.nr, ps.nth, ps.nz)
    return eafp.cg_3d_mpi(*[ps.padded(n) for n in COEF_NAMES], ps.Su, phi, m, max_iter, tol)


def check_halos(phi_out, phi_in, nth):
    # r/z halos untouched (Fortran i in {-1,0,nr+1,nr+2} -> numpy 0,1,-2,-1)
    for sl in ((slice(0, 2), slice(2, -2), slice(None)), (slice(-2, None), slice(2, -2), slice(None)),
               (slice(2, -2), slice(2, -2), slice(0, 2)), (slice(2, -2), slice(2, -2), slice(-2, None))):
        np.testing.assert_array_equal(phi_out[sl], phi_in[sl])
    # theta halos: x(:,0,:)=x(:,nth,:), x(:,-1,:)=x(:,nth-1,:), x(:,nth+1,:)=x(:,1,:), x(:,nth+2,:)=x(:,2,:)
    J = lambda j: j + 1
    np.testing.assert_array_equal(phi_out[:, J(0), :], phi_out[:, J(nth), :])
    np.testing.assert_array_equal(phi_out[:, J(-1), :], phi_out[:, J(nth - 1), :])
    np.testing.assert_array_equal(phi_out[:, J(nth + 1), :], phi_out[:, J(1), :])
    np.testing.assert_array_equal(phi_out[:, J(nth + 2), :], phi_out[:, J(2), :])


@pytest.mark.parametrize("config,shape", [(1, None), (2, (20, 36, 28)), (3, None), (5, (24, 48, 30))])
def test_drop_in_matches_oracle(config, shape):
    ps = systems.make_system(config, 0, shape)
    print(f"case config={config} shape={ps.shape}", flush=True)
    phi_in = np.zeros((ps.nr + 4, ps.nth + 4, ps.nz + 4), order="F")
    # poison r/z halos with finite values: wall links are zero, so they must
    # not change the answer and must come back untouched
    rng = np.random.default_rng(7)
    for sl in ((slice(0, 2), slice(None), slice(None)), (slice(-2, None), slice(None), slice(None)),
               (slice(None), slice(None), slice(0, 2)), (slice(None), slice(None), slice(-2, None))):
        phi_in[sl] = rng.normal(size=phi_in[sl].shape)
    phi = phi_in.copy(order="F")
    res, nit = drop_in(ps, phi)
    phi_ref, res_ref, nit_ref, _ = oracle.cg(ps, phi_in, max_iter=100000, tol=1e-8)
    assert res < 1e-8 and res_ref < 1e-8
    assert abs(nit - nit_ref) <= max(2, nit_ref // 50), (nit, nit_ref)
    err = oracle.gauge_error(oracle.interior(phi), oracle.interior(phi_ref), ps.labels, ps.ncomp)
    assert err < 1e-6, err
    l2, l1 = oracle.true_residual(ps, phi)
    assert l2 < 1.5e-8, l2
    check_halos(phi, phi_in, ps.nth)


def test_fold_and_initial_phi():
    """Nonzero wall links, r/z halos, interior x0 and RHS: the first residual
    takes each wall link times its halo value (the fold vector), as the
    reference's first matvec does; the iterates, iteration count and halos
    match the oracle, and dropping the halos changes the answer."""
    from fold_case import fold_case
    ps, phi0 = fold_case()
    print(f"case fold shape={ps.shape}", flush=True)
    m = eafp.Mesh(*ps.shape)
    phi = phi0.copy(order="F")
    res, nit = eafp.cg_3d_mpi(*[ps.padded(n) for n in COEF_NAMES], ps.Su, phi, m, 100000, 1e-8)
    phi_ref, res_ref, nit_ref, _ = oracle.cg(ps, phi0, max_iter=100000, tol=1e-8)
    print(f"  gpu n_iter={nit} res={res:.2e} | oracle n_iter={nit_ref} res={res_ref:.2e}", flush=True)
    assert res < 1e-8 and abs(nit - nit_ref) <= max(2, nit_ref // 50)
    x, xr = oracle.interior(phi), oracle.interior(phi_ref)
    assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < 1e-7
    l2, _ = oracle.true_residual(ps, phi)
    assert l2 < 1.5e-8
    check_halos(phi, phi0, ps.nth)
    # the same solve with zero r/z halos converges to another answer
    phi2 = phi0.copy(order="F")
    for sl in ((slice(0, 2),), (slice(-2, None),), (slice(None), slice(None), slice(0, 2)),
               (slice(None), slice(None), slice(-2, None))):
        phi2[sl] = 0.0
    eafp.cg_3d_mpi(*[ps.padded(n) for n in COEF_NAMES], ps.Su, phi2, m, 100000, 1e-8)
    assert np.linalg.norm(oracle.interior(phi2) - x) > 1e-3 * np.linalg.norm(x)
    # and max_iter = 3 stops on the oracle's third iterate
    phi3 = phi0.copy(order="F")
    eafp.cg_3d_mpi(*[ps.padded(n) for n in COEF_NAMES], ps.Su, phi3, m, 3, 1e-8)
    phi3_ref, _, _, _ = oracle.cg(ps, phi0, max_iter=3, tol=1e-8)
    np.testing.assert_allclose(oracle.interior(phi3), oracle.interior(phi3_ref), rtol=0,
                               atol=1e-9 * np.abs(oracle.interior(phi3_ref)).max())


def test_production_size_true_residual():
    """60x120x84 (604,800 cells): size-independent checks at full size."""
    ps = systems.make_system(2)
    print(f"case config=2 shape={ps.shape}", flush=True)
    phi = np.zeros((ps.nr + 4, ps.nth + 4, ps.nz + 4), order="F")
    res, nit = drop_in(ps, phi)
    assert res < 1e-8
    # reference FP64 Jacobi-PCG needs 1306 iterations on this system (oracle run, CPU)
    assert abs(nit - 1306) <= 30, nit
    A = oracle.csr(ps)
    x = oracle.interior(phi).ravel()
    b = ps.interior("Su").ravel()
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) < 1.5e-8
    check_halos(phi, np.zeros_like(phi), ps.nth)


def test_zero_rhs_early_return():
    ps = systems.make_system(1)
    ps.Su[...] = 0.0
    phi_in = np.zeros((ps.nr + 4, ps.nth + 4, ps.nz + 4), order="F")
    phi_in[2:-2, 2:-2, 2:-2] = 3.0   # interior left as is on the early return
    phi = phi_in.copy(order="F")
    res, nit = drop_in(ps, phi)
    assert res == 0.0 and nit == 0
    np.testing.assert_array_equal(phi[2:-2, 2:-2, 2:-2], phi_in[2:-2, 2:-2, 2:-2])
    check_halos(phi, phi_in, ps.nth)


@pytest.mark.parametrize("max_iter", [1, 2, 7])
def test_max_iter_cap(max_iter):
    ps = systems.make_system(1)
    phi = np.zeros((ps.nr + 4, ps.nth + 4, ps.nz + 4), order="F")
    res, nit = drop_in(ps, phi, max_iter=max_iter)
    phi_ref, res_ref, nit_ref, _ = oracle.cg(ps, max_iter=max_iter)
    assert nit == nit_ref == max_iter
    assert abs(res - res_ref) <= 1e-10 * res_ref
    np.testing.assert_allclose(phi[2:-2, 2:-2, 2:-2], phi_ref[2:-2, 2:-2, 2:-2], rtol=1e-9,
                               atol=1e-12 * np.abs(phi_ref).max())


def test_breakdown_exit():
    """aP = 0 and no links: p.q = 0 on the first iteration -> silent exit, n_iter = 0."""
    ps = systems.make_system(1)
    for n in COEF_NAMES:
        ps.coef[n][...] = 0.0
    phi = np.zeros((ps.nr + 4, ps.nth + 4, ps.nz + 4), order="F")
    res, nit = drop_in(ps, phi)
    phi_ref, res_ref, nit_ref, _ = oracle.cg(ps)
    assert nit == nit_ref == 0
    assert res == res_ref
    np.testing.assert_array_equal(phi[2:-2, 2:-2, 2:-2], 0.0)


def test_nan_input_propagates():
    """NaN in Su: the residual is NaN and the loop runs to max_iter, as in the reference."""
    ps = systems.make_system(1)
    ps.Su[5, 6, 7] = np.nan
    phi = np.zeros((ps.nr + 4, ps.nth + 4, ps.nz + 4), order="F")
    res, nit = drop_in(ps, phi, max_iter=5)
    _, res_ref, nit_ref, _ = oracle.cg(ps, max_iter=5)
    assert math.isnan(res) and math.isnan(res_ref)
    assert nit == nit_ref == 5


def test_parallel_mesh_refused():
    ps = systems.make_system(1)
    phi = np.zeros((ps.nr + 4, ps.nth + 4, ps.nz + 4), order="F")
    m = eafp.Mesh(ps.nr, ps.nth, ps.nz, is_parallel=True)
    res, nit = eafp.cg_3d_mpi(*[ps.padded(n) for n in COEF_NAMES], ps.Su, phi, m, 100, 1e-8)
    assert math.isnan(res) and nit == 0
    assert eafp.solver.last_status == eafp._lib.EAFP_ERR_UNSUPPORTED


def test_batch_fp64_ragged():
    """Uneven meshes in one launch sequence, per-system masks (workload 3 shapes)."""
    syss = [systems.make_system(3, s) for s in range(6)]
    syss.append(systems.make_system(1))
    z = systems.make_system(1, 1)
    z.Su[...] = 0.0                      # one system returns at once
    syss.append(z)
    with eafp.Batch.from_systems(syss) as b:
        st = b.solve_fp64(max_iter=100000, tol=1e-8)
        for s, ps in enumerate(syss):
            print(f"case batch system={s} shape={ps.shape}", flush=True)
            phi = b.download(s)
            phi_ref, res_ref, nit_ref, _ = oracle.cg(ps)
            if s == len(syss) - 1:
                assert st[s]["status"] == "zero_rhs" and st[s]["n_iter"] == 0
                continue
            assert st[s]["status"] == "converged"
            assert abs(st[s]["n_iter"] - nit_ref) <= max(2, nit_ref // 50)
            err = oracle.gauge_error(oracle.interior(phi), oracle.interior(phi_ref), ps.labels, ps.ncomp)
            assert err < 1e-6
        # a second solve of the same batch restarts from phi0 and repeats exactly
        st2 = b.solve_fp64(max_iter=100000, tol=1e-8)
        assert [x["n_iter"] for x in st2] == [x["n_iter"] for x in st]
        assert [x["residual"] for x in st2] == [x["residual"] for x in st]


@pytest.mark.parametrize("cpt", [1, 2, 4, 8])
def test_tilings_agree(cpt):
    """Every launch shape the autotuner may pick gives the same answer
    (only the partial-sum tree changes)."""
    syss = [systems.make_system(3, s) for s in range(3)]
    with eafp.Batch.from_systems(syss) as b:
        b.set_tiling(cpt, 0)
        st = b.solve_fp64(100000, 1e-8)
        for s, ps in enumerate(syss):
            phi = b.download(s)
            phi_ref, _, nit_ref, _ = oracle.cg(ps)
            assert st[s]["status"] == "converged"
            assert abs(st[s]["n_iter"] - nit_ref) <= max(2, nit_ref // 50)
            assert oracle.gauge_error(oracle.interior(phi), oracle.interior(phi_ref), ps.labels, ps.ncomp) < 1e-6


def test_autotune_keeps_results():
    ps = systems.make_system(2, 0, (20, 36, 28))
    with eafp.Batch.from_systems([ps]) as b:
        st0 = b.solve_fp64(100000, 1e-8)[0]
        best = b.autotune(eafp.Batch.SOLVER_FP64, reps=3)
        assert best["cells_per_thread"] in (1, 2, 4, 8) and best["ms_per_iter"] > 0
        info = b.info()
        assert info["tile_cells"] == 256 * best["cells_per_thread"] and info["grid"] == best["blocks"]
        st1 = b.solve_fp64(100000, 1e-8)[0]
        assert st1["status"] == "converged" and abs(st1["n_iter"] - st0["n_iter"]) <= 2
