"""The Fortran drop-in (mod_eafp_solver::cg_3d_mpi, ISO_C_BINDING to
eafp_cg_solve) compiles against libeafp.so, keeps the reference interface,
and maps library failures to residual = NaN (src/mod_pressure_3d.f90:322).
Runs on the CPU: the parallel-mesh refusal happens before any GPU call."""
import math
import os

import numpy as np
import pytest

from eafp import systems
from fortran_io import DRIVER, run_driver


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
