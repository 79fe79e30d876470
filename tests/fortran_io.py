"""Runs the Fortran drop-in drivers (hornofusion-full3d_amd/fortran/build/cg_driver,
tdma_driver and step_driver) on one case: raw stream files in the drivers'
documented layout."""
import os
import subprocess
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "hornofusion-full3d_amd", "fortran", "build", "cg_driver")
TDMA_DRIVER = os.path.join(ROOT, "hornofusion-full3d_amd", "fortran", "build", "tdma_driver")
STEP_DRIVER = os.path.join(ROOT, "hornofusion-full3d_amd", "fortran", "build", "step_driver")
STEP_FIELDS = ("vol", "Ar", "Ath", "Az", "alpha_l", "ur_l", "uth_l", "uz_l", "aPr_l", "aPth_l", "aPz_l", "rho_l", "T_l",
               "alpha_g", "ur_g", "uth_g", "uz_g", "aPr_g", "aPth_g", "aPz_g", "rho_g", "T_old_g", "p")
STEP_OUT = ("pp", "Fc_r", "Fc_th", "Fc_z", "lk_r", "lk_th", "lk_z", "ur_l", "uth_l", "uz_l", "ur_g", "uth_g", "uz_g", "p")
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


def run_step_driver(F, cfg, shape, alpha_p, max_iter, tol, is_parallel=False, solver=0, precond=0, timeout=120):
    """eafp_pressure_step from Fortran on the fields F (names of
    eafp._lib.PP_FIELDS; absent masks are not passed).  Returns (status,
    n_iter, residual, cg_res, out) with out mapping STEP_OUT to padded arrays."""
    nr, nth, nz = shape
    has = [int(n in F) for n in ("liq_cont", "pv_active", "gas_reach", "liq_room")]
    hdr = [nr, nth, nz, max_iter, int(is_parallel), int(cfg["gas_in_poisson"]), int(cfg["solve_multiphase"]),
           int(cfg["gas_compressibility"])] + has + [solver, precond]
    pad = (nr + 4, nth + 4, nz + 4)
    with tempfile.TemporaryDirectory() as d:
        fin = os.path.join(d, "in.bin")
        fout = os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            np.array(hdr, dtype=np.int32).tofile(f)
            np.array([cfg["dt"], cfg["rho_gas"], cfg["T_ambient"], cfg["beta_expansion"], alpha_p, tol],
                     dtype=np.float64).tofile(f)
            for n in ("r", "theta", "z"):
                np.asarray(F[n], dtype=np.float64).tofile(f)
            for n in STEP_FIELDS:
                np.asfortranarray(F[n], dtype=np.float64).ravel(order="F").tofile(f)
            np.asfortranarray(F["cell_type"], dtype=np.int32).ravel(order="F").tofile(f)
            np.asfortranarray(F["outlet"], dtype=np.int32).ravel(order="F").tofile(f)
            for n in ("liq_cont", "pv_active", "gas_reach"):
                if n in F:
                    np.asfortranarray(F[n], dtype=np.int32).ravel(order="F").tofile(f)
            if "liq_room" in F:
                np.asfortranarray(F["liq_room"], dtype=np.float64).ravel(order="F").tofile(f)
        r = subprocess.run([STEP_DRIVER, fin, fout], timeout=timeout, capture_output=True, text=True)
        assert r.returncode == 0, r.stdout + r.stderr
        raw = open(fout, "rb").read()
    status, n_iter = (int(v) for v in np.frombuffer(raw[:8], dtype=np.int32))
    residual, cg_res = (float(v) for v in np.frombuffer(raw[8:24], dtype=np.float64))
    ncell = int(np.prod(pad))
    out, pos = {}, 24
    for n in STEP_OUT:
        dt = np.int32 if n.startswith("lk") else np.float64
        nb = ncell * np.dtype(dt).itemsize
        out[n] = np.frombuffer(raw[pos:pos + nb], dtype=dt).reshape(pad, order="F").copy(order="F")
        pos += nb
    assert pos == len(raw)
    return status, n_iter, residual, cg_res, out
