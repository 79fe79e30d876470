"""Runs the Fortran drop-in drivers (hornofusion-full3d_amd/fortran/build/cg_driver
and tdma_driver) on one system: raw stream files in the drivers' documented layout."""
import os
import subprocess
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "hornofusion-full3d_amd", "fortran", "build", "cg_driver")
TDMA_DRIVER = os.path.join(ROOT, "hornofusion-full3d_amd", "fortran", "build", "tdma_driver")
NAMES = ("aW", "aE", "aS", "aN", "aB", "aT", "aP")


def run_driver(ps, phi, max_iter, tol, is_parallel=False, timeout=120):
    nr, nth, nz = ps.shape
    with tempfile.TemporaryDirectory() as d:
        fin = os.path.join(d, "in.bin")
        fout = os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            np.array([nr, nth, nz, max_iter, int(is_parallel)], dtype=np.int32).tofile(f)
            np.array([tol], dtype=np.float64).tofile(f)
            for n in NAMES:
                np.asfortranarray(ps.padded(n)).ravel(order="F").tofile(f)
            np.asfortranarray(ps.Su).ravel(order="F").tofile(f)
            np.asfortranarray(phi).ravel(order="F").tofile(f)
        subprocess.run([DRIVER, fin, fout], check=True, timeout=timeout)
        raw = open(fout, "rb").read()
    residual = np.frombuffer(raw[:8], dtype=np.float64)[0]
    n_iter, status = np.frombuffer(raw[8:16], dtype=np.int32)
    phi_out = np.frombuffer(raw[16:], dtype=np.float64).reshape((nr + 4, nth + 4, nz + 4), order="F").copy(order="F")
    return float(residual), int(n_iter), int(status), phi_out


def run_tdma_driver(coef, phi, shape, n_sweep, is_parallel=False, timeout=120):
    """tdma_3d_mpi from Fortran; coef maps aW..aT, aP, Su to padded arrays.
    Returns (status, phi_out), or (None, None) when the driver stopped with
    an error (error stop)."""
    nr, nth, nz = shape
    with tempfile.TemporaryDirectory() as d:
        fin = os.path.join(d, "in.bin")
        fout = os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            np.array([nr, nth, nz, n_sweep, int(is_parallel)], dtype=np.int32).tofile(f)
            for n in NAMES + ("Su",):
                np.asfortranarray(coef[n]).ravel(order="F").tofile(f)
            np.asfortranarray(phi).ravel(order="F").tofile(f)
        r = subprocess.run([TDMA_DRIVER, fin, fout], timeout=timeout, capture_output=True, text=True)
        if r.returncode != 0:
            return None, None
        raw = open(fout, "rb").read()
    status = int(np.frombuffer(raw[:4], dtype=np.int32)[0])
    phi_out = np.frombuffer(raw[4:], dtype=np.float64).reshape((nr + 4, nth + 4, nz + 4), order="F").copy(order="F")
    return status, phi_out
