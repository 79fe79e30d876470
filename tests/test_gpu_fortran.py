"""GPU parity of the Fortran drop-in: cg_3d_mpi called from Fortran exactly
as solve_pressure_correction calls it (src/mod_pressure_3d.f90:316-318),
against the oracle restatement of the reference CG."""
import numpy as np
import pytest

import oracle
from eafp import systems
from fortran_io import run_driver

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
