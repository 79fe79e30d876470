"""TEST INFRASTRUCTURE ONLY: the FP64 CPU oracle of EAF3D's p' CG solve.

Only ``tests/``, the smoke check and the benchmark's CPU-baseline leg import
this package.  It is the checker, never the thing measured or shipped.

Contents
--------
* ctypes bindings to ``liboracle.so`` (``cg_oracle.c``): a plain-C FP64
  restatement of ``cg_3d_mpi`` (``src/mod_solver_3d.f90:485-620``), its
  ``matvec`` (``:588-605``), ``dot`` (``:607-618``) and the serial theta halo
  copy (``src/mod_mpi_topology.f90:303-310``), on the reference's padded
  Fortran layout (bounds ``(-1:n+2)``, r fastest).
* ``pp_assemble`` / ``pp_correct``: ctypes bindings to ``pp_assemble.c``,
  the restatement of the p' system assembly of ``solve_pressure_correction``
  (``src/mod_pressure_3d.f90:171-306``, ``src/mod_boundary_3d.f90:180-277``)
  and of its post-solve steps (``:324-397``, ``:689-746``).
* ``tdma`` / ``tdma_3d``: ctypes bindings to ``tdma.c``, the restatement of
  the Thomas algorithm and of the line-by-line smoother ``tdma_3d_mpi``
  (``src/mod_solver_3d.f90:22-46``, ``:252-366``).
* numpy helpers that are independent of the C code and are used to pin it:
  a scipy CSR operator built from the interior coefficient arrays, the
  per-component zero-mean projection, and gauge-fixed error metrics.

Pinning (tests/test_oracle.py): scipy ``spsolve`` on a gauge-fixed system,
manufactured solutions ``b = A x*``, the closed form ``A 1 = aP - sum(a_nb)``,
theta-roll equivariance of the operator, ``math.fsum`` dots, and the CG
invariants (identity rows stay 0, ``b = 0`` returns at once, breakdown
exit, ``n_iter <= max_iter``).  The reference's exact gfortran residual
history is not available (no Fortran MPI toolchain in this image), so the
iteration-by-iteration history is "parity unpinned" beyond these checks.
"""
from __future__ import annotations

import contextlib
import ctypes
import math
import os
import subprocess
import tempfile

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_OMP_PATH = os.path.join(_HERE, "liboracle_omp.so")
_lib = None
_omp = None

_dp = ctypes.POINTER(ctypes.c_double)


def build(force: bool = False) -> str:
    """Compile ``liboracle.so`` and ``liboracle_omp.so`` with the committed
    Makefile (make rebuilds whatever is older than its sources)."""
    subprocess.run(["make", "-B" if force else "-s", "-C", _HERE], check=True, stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        build()
        _lib = _bind(ctypes.CDLL(_LIB_PATH))
    return _lib


def _bind(L):
    """Declares the argument and result types of every oracle function."""
    i = ctypes.c_int
    L.oracle_cg.argtypes = [i, i, i] + [_dp] * 8 + [_dp, i, ctypes.c_double, _dp,
                                                     ctypes.POINTER(ctypes.c_int), _dp]
    L.oracle_cg.restype = ctypes.c_int
    L.oracle_matvec.argtypes = [i, i, i] + [_dp] * 7 + [_dp, _dp]
    L.oracle_matvec.restype = None
    L.oracle_theta_halo.argtypes = [i, i, i, _dp]
    L.oracle_theta_halo.restype = None
    L.oracle_dot.argtypes = [i, i, i, _dp, _dp]
    L.oracle_dot.restype = ctypes.c_double
    L.oracle_true_residual.argtypes = [i, i, i] + [_dp] * 9 + [_dp, _dp]
    L.oracle_true_residual.restype = None
    L.oracle_pp_assemble.argtypes = [i, i, i, ctypes.c_void_p, ctypes.c_void_p] + [_dp] * 8 + [_dp, _dp]
    L.oracle_pp_assemble.restype = ctypes.c_int
    L.oracle_pp_assemble_ex.argtypes = [i, i, i, ctypes.c_void_p, ctypes.c_void_p] + [_dp] * 8 + \
        [_dp, _dp, ctypes.c_void_p]
    L.oracle_pp_assemble_ex.restype = ctypes.c_int
    ip = ctypes.POINTER(ctypes.c_int)
    L.oracle_pp_correct.argtypes = [i, i, i, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, _dp,
                                    ctypes.c_void_p, _dp, _dp, _dp, ip, ip, ip, ctypes.POINTER(_dp), _dp]
    L.oracle_pp_correct.restype = ctypes.c_int
    L.oracle_tdma.argtypes = [_dp] * 5 + [i]
    L.oracle_tdma.restype = ctypes.c_int
    L.oracle_tdma_3d.argtypes = [i, i, i] + [_dp] * 9 + [i]
    L.oracle_tdma_3d.restype = ctypes.c_int
    return L


_PLANT_DIR = None


@contextlib.contextmanager
def planted(n: int):
    """Within the block, every oracle call goes to the C sources rebuilt
    with ``-DORACLE_PLANT=n``: mistake ``n`` planted (the hooks are listed in
    ``pp_assemble.c`` and ``cg_oracle.c``).  The pinning tests use it to show
    that each of their checks fails on the mistake it is meant to catch."""
    global _lib, _PLANT_DIR
    if _PLANT_DIR is None:
        _PLANT_DIR = tempfile.mkdtemp(prefix="oracle_plant_")
    path = os.path.join(_PLANT_DIR, f"liboracle_plant{int(n)}.so")
    if not os.path.exists(path):
        srcs = [os.path.join(_HERE, f) for f in ("cg_oracle.c", "pp_assemble.c", "tdma.c")]
        subprocess.run(["gcc", "-O2", "-std=c11", "-fPIC", "-ffp-contract=off", "-shared",
                        f"-DORACLE_PLANT={int(n)}", "-o", path] + srcs + ["-lm"], check=True)
    saved = lib()
    _lib = _bind(ctypes.CDLL(path))
    try:
        yield
    finally:
        _lib = saved


def lib_omp():
    """``liboracle_omp.so``: the same CG with OpenMP loops (``cg_omp.c``)."""
    global _omp
    if _omp is None:
        if not os.path.exists(_OMP_PATH):
            build()
        L = ctypes.CDLL(_OMP_PATH)
        i = ctypes.c_int
        L.oracle_cg_omp.argtypes = [i, i, i] + [_dp] * 8 + [_dp, i, ctypes.c_double, _dp,
                                                             ctypes.POINTER(ctypes.c_int)]
        L.oracle_cg_omp.restype = ctypes.c_int
        _omp = L
    return _omp


PP_FIELDS = ("r", "theta", "z", "vol", "Ar", "Ath", "Az", "cell_type", "outlet",
             "alpha_l", "ur_l", "uth_l", "uz_l", "aPr_l", "aPth_l", "aPz_l", "rho_l", "T_l",
             "alpha_g", "ur_g", "uth_g", "uz_g", "aPr_g", "aPth_g", "aPz_g", "rho_g", "T_old_g", "p",
             "liq_cont", "pv_active", "gas_reach", "liq_room")


class PPFields(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in PP_FIELDS]


class PPConfig(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_double), ("rho_gas", ctypes.c_double), ("T_ambient", ctypes.c_double),
                ("beta_expansion", ctypes.c_double), ("gas_in_poisson", ctypes.c_int),
                ("solve_multiphase", ctypes.c_int), ("gas_compressibility", ctypes.c_int), ("pad", ctypes.c_int)]


def pp_structs(F, cfg):
    """ctypes (PPFields, PPConfig) over the arrays of ``F`` (absent masks: NULL)."""
    keep = []
    vals = {}
    for n in PP_FIELDS:
        a = F.get(n)
        if a is None:
            vals[n] = None
            continue
        want = np.int32 if n in ("cell_type", "outlet", "liq_cont", "pv_active", "gas_reach") else np.float64
        a = np.asfortranarray(a, dtype=want)
        keep.append(a)
        vals[n] = a.ctypes.data
    f = PPFields(**vals)
    c = PPConfig(float(cfg["dt"]), float(cfg["rho_gas"]), float(cfg["T_ambient"]), float(cfg["beta_expansion"]),
                 int(cfg["gas_in_poisson"]), int(cfg["solve_multiphase"]), int(cfg["gas_compressibility"]), 0)
    return f, c, keep


EXPORTS = ("Fs_r", "Fs_th", "Fs_z", "ac_r", "ac_th", "ac_z", "lk_r", "lk_th", "lk_z")


class PPExports(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in EXPORTS]


def _padded_zeros(shape, dtype=np.float64):
    nr, nth, nz = shape
    return np.zeros((nr + 4, nth + 4, nz + 4), dtype=dtype, order="F")


def pp_assemble(F, cfg, shape, exports=False):
    """The restated p' assembly.  Returns (coef dict with aW..aP and Su as
    padded arrays, residual, flux_ref), plus the dict of liquid face exports
    (Fs_*, ac_*, lk_*) when ``exports``."""
    f, c, keep = pp_structs(F, cfg)
    outs = {n: _padded_zeros(shape) for n in ("aW", "aE", "aS", "aN", "aB", "aT", "aP", "Su")}
    ex = {n: _padded_zeros(shape, np.int32 if n.startswith("lk") else np.float64) for n in EXPORTS} if exports else None
    exs = PPExports(**{n: ex[n].ctypes.data for n in EXPORTS}) if exports else None
    res = ctypes.c_double()
    flux = ctypes.c_double()
    rc = lib().oracle_pp_assemble_ex(*shape, ctypes.byref(f), ctypes.byref(c),
                                     *[outs[n].ctypes.data_as(_dp) for n in ("aW", "aE", "aS", "aN", "aB", "aT", "aP", "Su")],
                                     ctypes.byref(res), ctypes.byref(flux), ctypes.byref(exs) if exports else None)
    del keep
    if rc != 0:
        raise MemoryError("oracle_pp_assemble could not allocate its work arrays")
    if exports:
        return outs, res.value, flux.value, ex
    return outs, res.value, flux.value


VEL = ("ur_l", "uth_l", "uz_l", "ur_g", "uth_g", "uz_g")


def pp_correct(F, cfg, shape, alpha_p, pp, ex):
    """The restated post-solve consumers.  Returns a dict with the
    conservative fluxes Fc_r/th/z, their link flags lk_r/th/z, the corrected
    velocities (names of VEL) and the updated p; F is not modified."""
    f, c, keep = pp_structs(F, cfg)
    exs = PPExports(**{n: ex[n].ctypes.data for n in EXPORTS})
    out = {n: _padded_zeros(shape) for n in ("Fc_r", "Fc_th", "Fc_z")}
    out.update({n: _padded_zeros(shape, np.int32) for n in ("lk_r", "lk_th", "lk_z")})
    for n in VEL + ("p",):
        out[n] = np.array(F[n], dtype=np.float64, order="F")
    vel = (_dp * 6)(*[out[n].ctypes.data_as(_dp) for n in VEL])
    ppa = np.asfortranarray(pp, dtype=np.float64)
    ip = ctypes.POINTER(ctypes.c_int)
    rc = lib().oracle_pp_correct(*shape, ctypes.byref(f), ctypes.byref(c), float(alpha_p), ppa.ctypes.data_as(_dp),
                                 ctypes.byref(exs), *[out[n].ctypes.data_as(_dp) for n in ("Fc_r", "Fc_th", "Fc_z")],
                                 *[out[n].ctypes.data_as(ip) for n in ("lk_r", "lk_th", "lk_z")], vel,
                                 out["p"].ctypes.data_as(_dp))
    del keep
    if rc != 0:
        raise MemoryError("oracle_pp_correct could not allocate its work arrays")
    return out


def tdma(a, b, c, d):
    """The restated Thomas algorithm (``tdma``, src/mod_solver_3d.f90:22-46):
    x with a(i) x(i-1) + b(i) x(i) + c(i) x(i+1) = d(i)."""
    arrs = [np.ascontiguousarray(v, dtype=np.float64) for v in (a, b, c, d)]
    n = arrs[0].size
    x = np.empty(n)
    if lib().oracle_tdma(*[v.ctypes.data_as(_dp) for v in arrs], x.ctypes.data_as(_dp), n) != 0:
        raise ValueError("oracle_tdma: n < 1 or no memory")
    return x


TDMA_COEF = ("aW", "aE", "aS", "aN", "aB", "aT", "aP", "Su")


def tdma_3d(coef, phi, shape, n_sweep):
    """The restated ``tdma_3d_mpi`` on a serial mesh: ``coef`` maps aW..aT,
    aP, Su to padded Fortran arrays; returns the updated copy of ``phi``
    (padded, every halo as the reference leaves it)."""
    nr, nth, nz = shape
    arrs = [np.asfortranarray(coef[n], dtype=np.float64) for n in TDMA_COEF]
    out = np.array(phi, dtype=np.float64, order="F")
    for v in arrs + [out]:
        assert v.shape == (nr + 4, nth + 4, nz + 4)
    rc = lib().oracle_tdma_3d(nr, nth, nz, *[v.ctypes.data_as(_dp) for v in arrs], out.ctypes.data_as(_dp),
                              int(n_sweep))
    if rc != 0:
        raise ValueError("oracle_tdma_3d: bad shape or no memory")
    return out


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["F_CONTIGUOUS"], "padded arrays are Fortran-ordered float64"
    return a.ctypes.data_as(_dp)


COEF = ("aW", "aE", "aS", "aN", "aB", "aT", "aP")


def cg(ps, phi=None, max_iter=100000, tol=1e-8, history=False):
    """Run the restated ``cg_3d_mpi`` on system ``ps`` (padded layout).

    Returns ``(phi, residual, n_iter, hist)``; ``phi`` is a new padded array
    (x0 = ``phi`` or zero, as the reference's caller zeroes ``sh%pp``).
    """
    nr, nth, nz = ps.shape
    phi = np.zeros((nr + 4, nth + 4, nz + 4), order="F") if phi is None else np.array(phi, order="F", dtype=np.float64)
    res = ctypes.c_double()
    nit = ctypes.c_int()
    hist = np.full(max_iter + 1, np.nan) if history else None
    rc = lib().oracle_cg(nr, nth, nz, *[_p(ps.padded(n)) for n in COEF], _p(ps.padded("Su")),
                         _p(phi), int(max_iter), float(tol), ctypes.byref(res), ctypes.byref(nit),
                         hist.ctypes.data_as(_dp) if history else None)
    if rc != 0:
        raise MemoryError("oracle_cg could not allocate its work arrays")
    if history:
        hist = hist[: nit.value + 1]
    return phi, res.value, nit.value, hist


def cg_omp(ps, phi=None, max_iter=100000, tol=1e-8):
    """The multi-core CPU baseline: ``cg`` with OpenMP loops (threads from
    ``OMP_NUM_THREADS``).  Returns ``(phi, residual, n_iter)``."""
    nr, nth, nz = ps.shape
    phi = np.zeros((nr + 4, nth + 4, nz + 4), order="F") if phi is None else np.array(phi, order="F", dtype=np.float64)
    res = ctypes.c_double()
    nit = ctypes.c_int()
    rc = lib_omp().oracle_cg_omp(nr, nth, nz, *[_p(ps.padded(n)) for n in COEF], _p(ps.padded("Su")),
                                 _p(phi), int(max_iter), float(tol), ctypes.byref(res), ctypes.byref(nit))
    if rc != 0:
        raise MemoryError("oracle_cg_omp could not allocate its work arrays")
    return phi, res.value, nit.value


def matvec(ps, x):
    """y = A x with the reference's halo copy of x (x is modified in place)."""
    nr, nth, nz = ps.shape
    y = np.zeros_like(x, order="F")
    lib().oracle_matvec(nr, nth, nz, *[_p(ps.padded(n)) for n in COEF], _p(x), _p(y))
    return y


def theta_halo(shape, x):
    lib().oracle_theta_halo(*shape, _p(x))
    return x


def dot(shape, a, b):
    return lib().oracle_dot(*shape, _p(a), _p(b))


def true_residual(ps, x):
    """(||b - A x||/||b||, sum|b - A x| / sum|b|) with x in padded layout."""
    nr, nth, nz = ps.shape
    l2 = ctypes.c_double()
    l1 = ctypes.c_double()
    lib().oracle_true_residual(nr, nth, nz, *[_p(ps.padded(n)) for n in COEF], _p(ps.padded("Su")),
                               _p(np.asfortranarray(x)), ctypes.byref(l2), ctypes.byref(l1))
    return l2.value, l1.value


# ---------------------------------------------------------------------------
# numpy / scipy helpers, independent of cg_oracle.c
# ---------------------------------------------------------------------------

def csr(ps):
    """scipy CSR matrix of the interior operator, cells ordered z fastest.

    Row (i,j,k) holds aP on the diagonal and -a_nb on the neighbour columns,
    with the theta neighbour wrapping periodically (the halo copy made
    explicit) and r/z links at the walls dropped (their halo partners are 0).
    """
    import scipy.sparse as sp
    nr, nth, nz = ps.shape
    n = nr * nth * nz
    idx = np.arange(n).reshape(nr, nth, nz)
    rows = [idx.ravel()]
    cols = [idx.ravel()]
    vals = [ps.interior("aP").ravel()]
    for name, shift, axis in (("aW", 1, 0), ("aE", -1, 0), ("aS", 1, 1), ("aN", -1, 1),
                              ("aB", 1, 2), ("aT", -1, 2)):
        a = ps.interior(name)
        nb = np.roll(idx, shift, axis=axis)
        m = a != 0.0
        if axis != 1:   # non-periodic axes: a wall link would read a zero halo
            edge = np.zeros_like(m)
            if shift == 1:
                edge[(slice(None),) * axis + (0,)] = True
            else:
                edge[(slice(None),) * axis + (-1,)] = True
            m = m & ~edge
        rows.append(idx[m])
        cols.append(nb[m])
        vals.append(-a[m])
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))


def project(v, labels, ncomp):
    """Per-component zero-mean projection (FP64), masked cells -> 0."""
    v = np.asarray(v, dtype=np.float64).reshape(labels.shape)
    out = np.zeros_like(v)
    act = labels >= 0
    lab = labels[act]
    sums = np.bincount(lab, weights=v[act], minlength=ncomp)
    cnt = np.bincount(lab, minlength=ncomp)
    out[act] = v[act] - (sums / np.maximum(cnt, 1))[lab]
    return out


def fsum_dot(a, b):
    """Correctly rounded dot product (math.fsum of exact products' roundings)."""
    return math.fsum((np.asarray(a, dtype=np.float64) * np.asarray(b, dtype=np.float64)).ravel().tolist())


def gauge_error(x, x_ref, labels, ncomp):
    """||P x - P x_ref|| / ||P x_ref|| with P the per-component projection."""
    px = project(x, labels, ncomp)
    pr = project(x_ref, labels, ncomp)
    den = np.linalg.norm(pr)
    return float(np.linalg.norm(px - pr) / den) if den > 0 else float(np.linalg.norm(px - pr))


def interior(padded):
    return np.ascontiguousarray(padded[2:-2, 2:-2, 2:-2])
