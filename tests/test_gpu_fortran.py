"""GPU parity of the Fortran drop-in: cg_3d_mpi called from Fortran exactly
as solve_pressure_correction calls it (src/mod_pressure_3d.f90:316-318),
against the oracle restatement of the reference CG."""
import numpy as np
import pytest

import oracle
from eafp import systems
from fortran_io import STEP_OUT, run_driver, run_step_driver

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("config,shape,max_iter", [(1, None, 100000), (2, (20, 36, 28), 100000), (1, None, 50)])
def test_fortran_drop_in(config, shape, max_iter):
    ps = systems.make_system(config, 0, shape)
    print(f"case fortran config={config} shape={ps.shape} max_iter={max_iter}", flush=True)
    phi = np.zeros((ps.nr + 4, ps.nth + 4, ps.nz + 4), order="F")
    res, nit, status, phi_out = run_driver(ps, phi, max_iter, 1e-8)
    phi_ref, res_ref, nit_ref, _ = oracle.cg(ps, max_iter=max_iter, tol=1e-8)
    assert status == 0
    if max_iter < nit_ref or nit_ref == max_iter:
        assert nit == max_iter and abs(res - res_ref) <= 1e-9 * res_ref
    else:
        assert res < 1e-8 and abs(nit - nit_ref) <= max(2, nit_ref // 50)
    err = oracle.gauge_error(oracle.interior(phi_out), oracle.interior(phi_ref), ps.labels, ps.ncomp)
    assert err < 1e-6, err
    # theta halos are the periodic copies (the final halo exchange, src/mod_solver_3d.f90:567)
    np.testing.assert_array_equal(phi_out[:, 1, :], phi_out[:, ps.nth + 1, :])
    np.testing.assert_array_equal(phi_out[:, ps.nth + 2, :], phi_out[:, 2, :])


def test_fortran_fold_and_initial_phi():
    """cg_3d_mpi from Fortran with nonzero wall links, r/z halos, interior
    x0 and RHS: the oracle's iterates and halo semantics."""
    from fold_case import fold_case
    ps, phi0 = fold_case()
    print(f"case fortran fold shape={ps.shape}", flush=True)
    res, nit, status, phi_out = run_driver(ps, phi0, 100000, 1e-8)
    phi_ref, res_ref, nit_ref, _ = oracle.cg(ps, phi0, max_iter=100000, tol=1e-8)
    assert status == 0 and res < 1e-8 and abs(nit - nit_ref) <= max(2, nit_ref // 50)
    x, xr = oracle.interior(phi_out), oracle.interior(phi_ref)
    assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < 1e-7
    # r/z halos untouched (the theta copy overwrites their theta-halo corners)
    np.testing.assert_array_equal(phi_out[:2, 2:-2], phi0[:2, 2:-2])
    np.testing.assert_array_equal(phi_out[-2:, 2:-2], phi0[-2:, 2:-2])
    np.testing.assert_array_equal(phi_out[2:-2, 2:-2, :2], phi0[2:-2, 2:-2, :2])
    np.testing.assert_array_equal(phi_out[2:-2, 2:-2, -2:], phi0[2:-2, 2:-2, -2:])
    np.testing.assert_array_equal(phi_out[:, 1, :], phi_out[:, ps.nth + 1, :])
    np.testing.assert_array_equal(phi_out[:, ps.nth + 3, :], phi_out[:, 3, :])


@pytest.mark.parametrize("shape,n_sweep", [((12, 24, 16), 3), ((60, 120, 84), 1)])
def test_fortran_tdma_drop_in(shape, n_sweep):
    """tdma_3d_mpi called from Fortran: phi bit-identical to the oracle's restatement."""
    from fortran_io import run_tdma_driver
    print(f"case fortran tdma shape={shape} n_sweep={n_sweep}", flush=True)
    rng = np.random.default_rng(sum(shape))
    pad = tuple(n + 4 for n in shape)
    coef = {n: np.asfortranarray(rng.uniform(0.1, 1.0, pad)) for n in ("aW", "aE", "aS", "aN", "aB", "aT")}
    coef["aP"] = np.asfortranarray(sum(coef.values()) * 1.1)
    coef["Su"] = np.asfortranarray(rng.normal(size=pad))
    phi = np.asfortranarray(rng.normal(size=pad))
    status, phi_out = run_tdma_driver(coef, phi, shape, n_sweep)
    assert status == 0
    want = oracle.tdma_3d(coef, phi, shape, n_sweep)
    assert np.count_nonzero(phi_out.view(np.uint64) != want.view(np.uint64)) == 0


def test_fortran_tdma_refuses_parallel_mesh():
    """A parallel mesh is refused and the program stops (the subroutine has no status)."""
    from fortran_io import run_tdma_driver
    shape = (4, 6, 5)
    pad = tuple(n + 4 for n in shape)
    coef = {n: np.ones(pad, order="F") for n in ("aW", "aE", "aS", "aN", "aB", "aT", "Su")}
    coef["aP"] = np.full(pad, 7.0, order="F")
    status, _ = run_tdma_driver(coef, np.zeros(pad, order="F"), shape, 1, is_parallel=True)
    assert status is None


@pytest.mark.parametrize("shape,kw,solver,precond", [((20, 36, 28), dict(seed=1, masks=True), 0, 0),
                                                     ((12, 24, 16), dict(seed=0), 0, 0),
                                                     ((16, 30, 24), dict(seed=2, two_phase=False), 1, 1),
                                                     ((60, 120, 84), dict(seed=5, masks=True), 0, 0)])
def test_fortran_pressure_step(shape, kw, solver, precond):
    """eafp_pressure_step from Fortran, velocities and p updated in place,
    bit for bit the C-ABI step (eafp_pp_step through ctypes) on the same
    fields; the FP64 form is itself bit-identical to the three separate
    entry points (tests/test_gpu_step.py)."""
    import eafp
    from eafp import fields
    print(f"case fortran pressure step shape={shape} {kw} solver={solver} precond={precond}", flush=True)
    F, cfg = fields.make_fields(*shape, **kw)
    st, nit, res, cg_res, out = run_step_driver(F, cfg, shape, 0.3, 50, 1e-8, solver=solver, precond=precond)
    assert st == 0
    ref, pp, res_ref, sref = eafp.pp_step(F, cfg, eafp.Mesh(*shape), 0.3, solver="mixed" if solver else "fp64",
                                          max_iter=50, tol=1e-8,
                                          mixed=dict(tol=1e-8, interval=32, max_steps=2000, precond=precond))
    assert res == res_ref and cg_res == sref["residual"] and nit == sref["n_iter"]
    np.testing.assert_array_equal(out["pp"], pp)
    for n in STEP_OUT[1:]:
        np.testing.assert_array_equal(out[n], ref[n], err_msg=n)
