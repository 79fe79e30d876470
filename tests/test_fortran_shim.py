"""The Fortran drop-in (mod_eafp_solver::cg_3d_mpi, ISO_C_BINDING to
eafp_cg_solve) compiles against libeafp.so, keeps the reference interface,
and maps library failures to residual = NaN (src/mod_pressure_3d.f90:322).
Runs on the CPU: the parallel-mesh refusal happens before any GPU call."""
import math
import os

import numpy as np
import pytest

from eafp import systems
from fortran_io import DRIVER, STEP_DRIVER, run_driver, run_step_driver


def test_driver_built():
    assert os.access(DRIVER, os.X_OK), "build the Fortran shim first (__graft_entry__.build())"


def test_parallel_mesh_gives_nan():
    ps = systems.make_system(1)
    phi = np.zeros((ps.nr + 4, ps.nth + 4, ps.nz + 4), order="F")
    phi[3, 4, 5] = 7.0
    res, nit, status, phi_out = run_driver(ps, phi, 100, 1e-8, is_parallel=True)
    assert math.isnan(res) and nit == 0
    assert status == -2                      # EAFP_ERR_UNSUPPORTED
    np.testing.assert_array_equal(phi_out, phi)


def test_step_driver_built():
    assert os.access(STEP_DRIVER, os.X_OK), "build the Fortran shim first (__graft_entry__.build())"


@pytest.mark.parametrize("masks", [True, False])
def test_pressure_step_binding_reaches_the_device_check(masks):
    """eafp_pressure_step builds the bind(C) structs from the module arrays:
    every required pointer arrives non-NULL, dt > 0 and the solver code is
    valid, so the library's checks pass and stop only at the device check
    here (EAFP_ERR_NO_DEVICE without a GPU; a GPU run returns 0).  On failure
    nothing is updated and the residuals are NaN."""
    from eafp import fields
    shape = (6, 8, 5)
    F, cfg = fields.make_fields(*shape, seed=0, masks=masks)
    st, nit, res, cg_res, out = run_step_driver(F, cfg, shape, 0.3, 50, 1e-8)
    assert st in (-4, 0), st
    if st == -4:
        assert math.isnan(res) and math.isnan(cg_res) and nit == 0
        for n in ("ur_l", "uz_g", "p"):
            np.testing.assert_array_equal(out[n], F[n])


def test_pressure_step_refusals():
    """A parallel mesh (EAFP_ERR_UNSUPPORTED), dt = 0 and an unknown solver
    code (EAFP_ERR_ARG) are refused before any device call."""
    from eafp import fields
    shape = (6, 8, 5)
    F, cfg = fields.make_fields(*shape, seed=1)
    st, nit, res, _, out = run_step_driver(F, cfg, shape, 0.3, 50, 1e-8, is_parallel=True)
    assert st == -2 and math.isnan(res) and nit == 0
    np.testing.assert_array_equal(out["p"], F["p"])
    st, _, res, _, _ = run_step_driver(F, dict(cfg, dt=0.0), shape, 0.3, 50, 1e-8)
    assert st == -1 and math.isnan(res)
    st, _, res, _, _ = run_step_driver(F, cfg, shape, 0.3, 50, 1e-8, solver=3)
    assert st == -1 and math.isnan(res)
