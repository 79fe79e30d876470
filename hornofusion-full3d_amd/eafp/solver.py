"""Python mirror of the reference's p' solver interface, on libeafp.so.

``cg_3d_mpi`` keeps the argument list and meaning of the Fortran subroutine
(``src/mod_solver_3d.f90:485-496``): seven coefficient arrays and ``Su`` in
the padded ``(-1:n+2)`` Fortran layout, ``phi`` updated in place, the mesh,
``max_iter`` and ``tol`` in; ``residual`` and ``n_iter`` out.  As in the
reference, failure is reported through a NaN residual (the library's status
is kept in ``last_status``).

``Batch`` holds one or more systems resident on the GPU for repeated solves:
the plain FP64 CG (the reference algorithm) and the mixed-precision solver
(FP32 inner CG with FP64 iterative refinement).
"""
from __future__ import annotations

import ctypes
import dataclasses
import math

import numpy as np

from . import _lib
from ._lib import Desc, MixedOpts, Stats, check

COEF_NAMES = ("aW", "aE", "aS", "aN", "aB", "aT", "aP")


@dataclasses.dataclass
class Mesh:
    """The fields of ``mesh_t`` that ``cg_3d_mpi`` reads (src/mod_types_3d.f90:143-193)."""
    nr: int
    ntheta: int
    nz: int
    is_parallel: bool = False
    periodic_theta: bool = True


def desc_of(m) -> Desc:
    return Desc(int(m.nr), int(m.ntheta), int(m.nz), 2, -1, int(bool(m.is_parallel)),
                int(bool(getattr(m, "periodic_theta", True))))


def _padded(a, shape, name, writable=False):
    nr, nth, nz = shape
    if not isinstance(a, np.ndarray) or a.dtype != np.float64:
        raise TypeError(f"{name} must be a float64 numpy array")
    if a.shape != (nr + 4, nth + 4, nz + 4) or not a.flags["F_CONTIGUOUS"]:
        raise ValueError(f"{name} must be Fortran-ordered with bounds (-1:n+2): shape {(nr + 4, nth + 4, nz + 4)}")
    if writable and not a.flags["WRITEABLE"]:
        raise ValueError(f"{name} must be writable")
    return a.ctypes.data


last_status = 0


def cg_3d_mpi(aW, aE, aS, aN, aB, aT, aP, Su, phi, m, max_iter, tol):
    """Drop-in for ``cg_3d_mpi``; returns ``(residual, n_iter)``, updates ``phi``.

    A nonzero library status (bad mesh descriptor, parallel mesh, HIP error)
    gives ``residual = nan`` and ``n_iter = 0``, the failure signal the
    reference's caller consumes (src/mod_pressure_3d.f90:322).
    """
    global last_status
    L = _lib.load()
    d = desc_of(m)
    shape = (m.nr, m.ntheta, m.nz)
    ptrs = [_padded(a, shape, n) for a, n in zip((aW, aE, aS, aN, aB, aT, aP, Su), COEF_NAMES + ("Su",))]
    ptrs.append(_padded(phi, shape, "phi", writable=True))
    res = ctypes.c_double(0.0)
    nit = ctypes.c_int(0)
    rc = L.eafp_cg_solve(ctypes.byref(d), *ptrs, int(max_iter), float(tol), ctypes.byref(res), ctypes.byref(nit))
    last_status = rc
    if rc != 0:
        return math.nan, 0
    return res.value, nit.value


def _field_shape(n, m):
    """The padded shape the library reads for field ``n`` on mesh ``m``: the
    coordinates r, theta, z have bounds (-1:n+2), the outlet mask is the top
    layer (nr+4, ntheta+4), every other field (nr+4, ntheta+4, nz+4)."""
    if n == "r":
        return (m.nr + 4,)
    if n == "theta":
        return (m.ntheta + 4,)
    if n == "z":
        return (m.nz + 4,)
    if n == "outlet":
        return (m.nr + 4, m.ntheta + 4)
    return (m.nr + 4, m.ntheta + 4, m.nz + 4)


def _checked(a, shape, name):
    """``a`` as the library's padded array, or ValueError when its shape is
    not ``shape`` (the library copies exactly that many values)."""
    if np.shape(a) != shape:
        raise ValueError(f"{name} must have the padded shape {shape}, not {np.shape(a)}")
    return a


def _pp_structs(F, cfg, m):
    """eafp_pp_fields / eafp_pp_config for the arrays of ``F`` (absent
    workspace masks are not valid; each array checked against its padded
    shape on mesh ``m``) and the config dict ``cfg``; ``keep`` holds the
    converted arrays alive."""
    keep = []
    vals = {}
    for n in _lib.PP_FIELDS:
        a = F.get(n)
        if a is None:
            vals[n] = None
            continue
        _checked(a, _field_shape(n, m), n)
        a = np.asfortranarray(a, dtype=np.int32 if n in _lib.PP_INT else np.float64)
        keep.append(a)
        vals[n] = a.ctypes.data
    f = _lib.PPFields(**vals)
    c = _lib.PPConfig(float(cfg["dt"]), float(cfg["rho_gas"]), float(cfg["T_ambient"]), float(cfg["beta_expansion"]),
                      int(cfg["gas_in_poisson"]), int(cfg["solve_multiphase"]), int(cfg["gas_compressibility"]), 0)
    return f, c, keep


def _padded_zeros(m, dtype=np.float64):
    return np.zeros((m.nr + 4, m.ntheta + 4, m.nz + 4), dtype=dtype, order="F")


def pp_assemble(F, cfg, m, exports=False):
    """The p' system assembled on the GPU (eafp_pp_assemble) from the fields
    of ``F`` (padded Fortran arrays, names as in ``_lib.PP_FIELDS``; absent
    workspace masks are not valid) and the config dict ``cfg``.  Returns
    ``(coef, residual, flux_ref)`` with ``coef`` mapping aW..aT, aP, Su to new
    padded arrays, plus the dict of the liquid's face exports (names of
    ``_lib.PP_EXPORTS``) when ``exports``."""
    L = _lib.load()
    f, c, keep = _pp_structs(F, cfg, m)
    names = ("aW", "aE", "aS", "aN", "aB", "aT", "aP", "Su")
    out = {n: _padded_zeros(m) for n in names}
    ex = {n: _padded_zeros(m, np.int32 if n.startswith("lk") else np.float64) for n in _lib.PP_EXPORTS} if exports else None
    exs = _lib.PPExports(**{n: ex[n].ctypes.data for n in _lib.PP_EXPORTS}) if exports else None
    res = ctypes.c_double()
    flux = ctypes.c_double()
    d = desc_of(m)
    check(L.eafp_pp_assemble(ctypes.byref(d), ctypes.byref(f), ctypes.byref(c), *[out[n].ctypes.data for n in names],
                             ctypes.byref(res), ctypes.byref(flux), ctypes.byref(exs) if exports else None, 0),
          "eafp_pp_assemble")
    del keep
    if exports:
        return out, res.value, flux.value, ex
    return out, res.value, flux.value


def pp_correct(F, cfg, m, alpha_p, pp, ex):
    """The post-solve consumers on the GPU (eafp_pp_correct): returns a dict
    with the conservative fluxes Fc_r/th/z, their link flags lk_r/th/z, the
    corrected velocities (ur_l .. uz_g) and the updated p, all new padded
    arrays; ``F`` is not modified.  ``pp`` is the solved correction (theta
    halos filled), ``ex`` the exports of ``pp_assemble(..., exports=True)``."""
    L = _lib.load()
    f, c, keep = _pp_structs(F, cfg, m)
    pshape = _field_shape("p", m)
    exa = {n: np.asfortranarray(_checked(ex[n], pshape, n), dtype=np.int32 if n.startswith("lk") else np.float64)
           for n in _lib.PP_EXPORTS}
    exs = _lib.PPExports(**{n: a.ctypes.data for n, a in exa.items()})
    out = {n: _padded_zeros(m, np.int32 if n.startswith("lk") else np.float64) for n in _lib.PP_UPDATE[:6]}
    for n in _lib.PP_UPDATE[6:]:
        out[n] = np.array(F[n], dtype=np.float64, order="F")
    io = _lib.PPUpdate(**{n: out[n].ctypes.data for n in _lib.PP_UPDATE})
    ppa = np.asfortranarray(_checked(pp, pshape, "pp"), dtype=np.float64)
    d = desc_of(m)
    check(L.eafp_pp_correct(ctypes.byref(d), ctypes.byref(f), ctypes.byref(c), float(alpha_p), ppa.ctypes.data,
                            ctypes.byref(exs), ctypes.byref(io), 0), "eafp_pp_correct")
    del keep
    return out


def pp_step(F, cfg, m, alpha_p, solver="fp64", max_iter=1000, tol=1e-8, mixed=None, in_place=False):
    """One whole p' step on the GPU (eafp_pp_step): assembly, the solve from
    pp = 0 and the post-solve steps, with the p' system kept in HBM.
    ``solver`` "fp64" is cg_3d_mpi's Jacobi-PCG (``max_iter``, ``tol``);
    "mixed" the mixed-precision solver with ``mixed`` (a dict of MixedOpts
    fields; tol 1e-8 and BF16 coefficients by default).  Returns ``(out,
    pp, residual, stats)``: ``out`` as ``pp_correct`` returns it, ``pp`` the
    solved correction, ``residual`` the continuity residual (NaN when the
    solve broke down) and ``stats`` the solve's outcome.  ``in_place``: the
    velocities and p of ``F`` (Fortran-ordered float64) are updated where
    they lie and returned in ``out``, as the reference updates liq%ur and
    sh%p, which are also the assembly's inputs."""
    L = _lib.load()
    f, c, keep = _pp_structs(F, cfg, m)
    out = {n: _padded_zeros(m, np.int32 if n.startswith("lk") else np.float64) for n in _lib.PP_UPDATE[:6]}
    for n in _lib.PP_UPDATE[6:]:
        if in_place:
            a = F[n]
            if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.f_contiguous):
                raise ValueError(f"in_place needs F[{n!r}] as a Fortran-ordered float64 array")
            out[n] = a
        else:
            out[n] = np.array(F[n], dtype=np.float64, order="F")
    io = _lib.PPUpdate(**{n: out[n].ctypes.data for n in _lib.PP_UPDATE})
    mo = dict(tol=1e-8, interval=32, max_steps=2000, coef_format=_lib.COEF_BF16, restart=0,
              precond=_lib.PRECOND_JACOBI)
    mo.update(mixed or {})
    if solver not in ("fp64", "mixed"):
        raise ValueError(f"solver must be 'fp64' or 'mixed', not {solver!r}")
    o = _lib.PPStepOpts(_lib.STEP_FP64 if solver == "fp64" else _lib.STEP_MIXED, int(max_iter), float(tol),
                        float(alpha_p), _lib.MixedOpts(**mo))
    pp = _padded_zeros(m)
    res = ctypes.c_double()
    st = _lib.Stats()
    d = desc_of(m)
    check(L.eafp_pp_step(ctypes.byref(d), ctypes.byref(f), ctypes.byref(c), ctypes.byref(o), pp.ctypes.data,
                         ctypes.byref(io), ctypes.byref(res), ctypes.byref(st), 0), "eafp_pp_step")
    del keep
    return out, pp, res.value, st.as_dict()


def tdma_3d_mpi(aW, aE, aS, aN, aB, aT, aP, Su, phi, m, n_sweep):
    """Mirror of ``tdma_3d_mpi(aW, aE, aS, aN, aB, aT, aP, Su, phi, m, n_sweep)``
    (src/mod_solver_3d.f90:252-366) on the GPU (eafp_tdma_solve): ``n_sweep``
    line-by-line sweeps, ``phi`` (padded Fortran float64) updated in place,
    halos included, exactly as the reference leaves it.  Raises on failure
    (the reference's subroutine has no status to report one)."""
    L = _lib.load()
    d = desc_of(m)
    shape = (m.nr, m.ntheta, m.nz)
    ptrs = [_padded(a, shape, n) for a, n in zip((aW, aE, aS, aN, aB, aT, aP, Su), COEF_NAMES + ("Su",))]
    ptrs.append(_padded(phi, shape, "phi", writable=True))
    check(L.eafp_tdma_solve(ctypes.byref(d), *ptrs, int(n_sweep), 0), "eafp_tdma_solve")


def tdma_batch(systems, meshes, n_sweep):
    """``n_sweep`` sweeps on several independent systems at once
    (eafp_tdma_solve_batch).  ``systems``: dicts with aW..aT, aP, Su and phi
    (padded float64, phi updated in place); ``meshes``: one Mesh each."""
    L = _lib.load()
    nsys = len(systems)
    descs = (_lib.Desc * nsys)(*[desc_of(m) for m in meshes])
    names = COEF_NAMES + ("Su", "phi")
    tab = (_lib.TdmaSystem * nsys)()
    for q, (sy, m) in enumerate(zip(systems, meshes)):
        shape = (m.nr, m.ntheta, m.nz)
        vals = {n: _padded(sy[n], shape, n, writable=(n == "phi")) for n in names}
        tab[q] = _lib.TdmaSystem(**vals)
    check(L.eafp_tdma_solve_batch(nsys, descs, tab, int(n_sweep), 0), "eafp_tdma_solve_batch")


def counters():
    """The library's process-wide activity counters (eafp_get_counters)."""
    L = _lib.load()
    c = _lib.Counters()
    check(L.eafp_get_counters(ctypes.byref(c)), "eafp_get_counters")
    return {n: int(getattr(c, n)) for n, _ in _lib.Counters._fields_}


def release():
    _lib.load().eafp_release()


class Batch:
    """Systems resident in HBM: upload once, solve many times."""

    def __init__(self, shapes):
        L = _lib.load()
        self._L = L
        self.shapes = [tuple(int(v) for v in s) for s in shapes]
        n = len(self.shapes)
        arr = (Desc * n)(*[desc_of(Mesh(*s)) for s in self.shapes])
        h = ctypes.c_void_p()
        check(L.eafp_batch_create(n, arr, ctypes.byref(h)), "eafp_batch_create")
        self._h = h

    @classmethod
    def from_systems(cls, systems):
        b = cls([ps.shape for ps in systems])
        for s, ps in enumerate(systems):
            b.upload(s, ps)
        return b

    def __len__(self):
        return len(self.shapes)

    def upload(self, s, ps, phi0=None):
        shape = self.shapes[s]
        ptrs = [_padded(ps.padded(n), shape, n) for n in COEF_NAMES + ("Su",)]
        p0 = _padded(phi0, shape, "phi0") if phi0 is not None else None
        check(self._L.eafp_batch_upload(self._h, s, *ptrs, p0, 0), "eafp_batch_upload")

    def upload_device(self, s, ptrs, phi0=None):
        """Uploads system s from device memory: ``ptrs`` are the device
        addresses of its padded aW, aE, aS, aN, aB, aT, aP and Su (packed
        where they lie); ``phi0`` a device address or None (zero)."""
        if len(ptrs) != 8:
            raise ValueError("upload_device needs the 8 arrays aW..aT, aP, Su")
        check(self._L.eafp_batch_upload(self._h, s, *[ctypes.c_void_p(int(p)) for p in ptrs],
                                        None if phi0 is None else ctypes.c_void_p(int(phi0)), 1),
              "eafp_batch_upload")

    def solve_fp64(self, max_iter=100000, tol=1e-8):
        st = (Stats * len(self))()
        check(self._L.eafp_batch_solve_fp64(self._h, int(max_iter), float(tol), st), "eafp_batch_solve_fp64")
        return [s.as_dict() for s in st]

    def solve_mixed(self, tol=1e-8, interval=32, max_steps=2000, coef_format=_lib.COEF_FP32, restart=False,
                    precond=_lib.PRECOND_JACOBI):
        """FP32 inner CG + FP64 refinement every `interval` inner iterations;
        ``precond`` the inner preconditioner (PRECOND_JACOBI or PRECOND_ZLINE)."""
        o = MixedOpts(float(tol), int(interval), int(max_steps), int(coef_format), int(bool(restart)), int(precond))
        st = (Stats * len(self))()
        check(self._L.eafp_batch_solve_mixed(self._h, ctypes.byref(o), st), "eafp_batch_solve_mixed")
        return [s.as_dict() for s in st]

    SOLVER_FP64, SOLVER_MIXED, SOLVER_MIXED_ZLINE = 0, 1, 2   # the last: mixed with the z-line preconditioner

    def set_tiling(self, cells_per_thread, blocks=0):
        check(self._L.eafp_batch_set_tiling(self._h, int(cells_per_thread), int(blocks)), "eafp_batch_set_tiling")

    def set_columns(self, r_splits):
        """Mixed stencil as a column march with r_splits r-segments per column (0: tile kernels)."""
        check(self._L.eafp_batch_set_columns(self._h, int(r_splits)), "eafp_batch_set_columns")

    def time_kernels(self, solver, coef_format=_lib.COEF_FP32, reps=20):
        """Average ms of one launch of each iteration kernel (spmv, update)."""
        a = ctypes.c_float()
        b = ctypes.c_float()
        check(self._L.eafp_batch_time_kernels(self._h, int(solver), int(coef_format), int(reps),
                                              ctypes.byref(a), ctypes.byref(b)), "eafp_batch_time_kernels")
        return a.value, b.value

    def autotune(self, solver, coef_format=_lib.COEF_FP32, reps=10):
        cpt = ctypes.c_int()
        blocks = ctypes.c_int()
        ms = ctypes.c_float()
        check(self._L.eafp_batch_autotune(self._h, int(solver), int(coef_format), int(reps), ctypes.byref(cpt),
                                          ctypes.byref(blocks), ctypes.byref(ms)), "eafp_batch_autotune")
        return {"cells_per_thread": cpt.value, "blocks": blocks.value, "ms_per_iter": ms.value}

    def components(self, s):
        """System s's singular components: (how many, their cell counts in
        order of first cell when there are at most 4, else [])."""
        n = ctypes.c_int()
        cells = (ctypes.c_int * 4)()
        check(self._L.eafp_batch_components(self._h, int(s), ctypes.byref(n), cells), "eafp_batch_components")
        return n.value, [cells[q] for q in range(n.value)] if n.value <= 4 else []

    def download(self, s, phi=None):
        nr, nth, nz = self.shapes[s]
        if phi is None:
            phi = np.zeros((nr + 4, nth + 4, nz + 4), order="F")
        check(self._L.eafp_batch_download(self._h, s, _padded(phi, self.shapes[s], "phi", True), 0),
              "eafp_batch_download")
        return phi

    def info(self):
        g = ctypes.c_int()
        t = ctypes.c_int()
        n = ctypes.c_long()
        check(self._L.eafp_batch_info(self._h, ctypes.byref(g), ctypes.byref(t), ctypes.byref(n)), "eafp_batch_info")
        c = ctypes.c_int()
        check(self._L.eafp_batch_get_columns(self._h, ctypes.byref(c)), "eafp_batch_get_columns")
        return {"grid": g.value, "tile_cells": t.value, "ncells": n.value, "column_splits": c.value}

    def close(self):
        if getattr(self, "_h", None):
            self._L.eafp_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
