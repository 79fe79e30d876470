"""ctypes binding of libeafp.so (the C-ABI declared in include/eafp.h).

The library is the only compute path: if it is missing this module raises,
it never falls back to a CPU implementation.
"""
from __future__ import annotations

import ctypes
import os

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG, "lib", "libeafp.so")
CSRC = os.path.join(_PKG, "csrc")

EAFP_OK = 0
EAFP_ERR_ARG = -1
EAFP_ERR_UNSUPPORTED = -2
EAFP_ERR_COMPONENTS = -3
EAFP_ERR_NO_DEVICE = -4
EAFP_ERR_STATE = -5
EAFP_ERR_SIZE = -6
NOALLOC_MAX_CELLS = 1048576       # EAFP_NOALLOC_MAX_CELLS

STOP_NAMES = {0: "running", 1: "converged", 2: "breakdown", 3: "max_iter", 4: "zero_rhs"}

COEF_FP32, COEF_BF16, COEF_FP16 = 0, 1, 2
PRECOND_JACOBI, PRECOND_ZLINE = 0, 1

# Every symbol declared in include/eafp.h.
EXPORTS = (
    "eafp_cg_solve", "eafp_release", "eafp_pp_assemble", "eafp_pp_assemble_noalloc", "eafp_pp_correct", "eafp_pp_step", "eafp_tdma_solve",
    "eafp_tdma_solve_batch", "eafp_batch_create", "eafp_batch_upload",
    "eafp_batch_solve_fp64", "eafp_batch_solve_mixed", "eafp_batch_download",
    "eafp_batch_info", "eafp_batch_set_tiling", "eafp_batch_set_columns", "eafp_batch_get_columns",
    "eafp_batch_time_kernels", "eafp_batch_autotune", "eafp_batch_components", "eafp_batch_destroy",
    "eafp_get_counters", "eafp_version",
)


PP_FIELDS = ("r", "theta", "z", "vol", "Ar", "Ath", "Az", "cell_type", "outlet",
             "alpha_l", "ur_l", "uth_l", "uz_l", "aPr_l", "aPth_l", "aPz_l", "rho_l", "T_l",
             "alpha_g", "ur_g", "uth_g", "uz_g", "aPr_g", "aPth_g", "aPz_g", "rho_g", "T_old_g", "p",
             "liq_cont", "pv_active", "gas_reach", "liq_room")
PP_INT = ("cell_type", "outlet", "liq_cont", "pv_active", "gas_reach")


class PPFields(ctypes.Structure):
    """eafp_pp_fields: one pointer per field, in the header's order."""
    _fields_ = [(n, ctypes.c_void_p) for n in PP_FIELDS]


PP_EXPORTS = ("Fs_r", "Fs_th", "Fs_z", "ac_r", "ac_th", "ac_z", "lk_r", "lk_th", "lk_z")
PP_UPDATE = ("Fc_r", "Fc_th", "Fc_z", "lk_r", "lk_th", "lk_z",
             "ur_l", "uth_l", "uz_l", "ur_g", "uth_g", "uz_g", "p")


class PPExports(ctypes.Structure):
    """eafp_pp_exports: the liquid's face terms (lk_* are int arrays)."""
    _fields_ = [(n, ctypes.c_void_p) for n in PP_EXPORTS]


class PPUpdate(ctypes.Structure):
    """eafp_pp_update: what the post-solve consumers write."""
    _fields_ = [(n, ctypes.c_void_p) for n in PP_UPDATE]


class TdmaSystem(ctypes.Structure):
    """eafp_tdma_system: aW..aT, aP, Su and phi of one system."""
    _fields_ = [(n, ctypes.c_void_p) for n in ("aW", "aE", "aS", "aN", "aB", "aT", "aP", "Su", "phi")]


class PPConfig(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_double), ("rho_gas", ctypes.c_double), ("T_ambient", ctypes.c_double),
                ("beta_expansion", ctypes.c_double), ("gas_in_poisson", ctypes.c_int),
                ("solve_multiphase", ctypes.c_int), ("gas_compressibility", ctypes.c_int), ("pad_", ctypes.c_int)]


class EafpError(RuntimeError):
    def __init__(self, rc, where):
        names = {EAFP_ERR_ARG: "invalid argument", EAFP_ERR_UNSUPPORTED: "unsupported mesh/descriptor",
                 EAFP_ERR_COMPONENTS: "too many singular components", EAFP_ERR_NO_DEVICE: "no HIP device",
                 EAFP_ERR_STATE: "invalid state", EAFP_ERR_SIZE: "mesh beyond the allocation-free scratch"}
        msg = names.get(rc, f"HIP error {rc}" if rc > 0 else f"error {rc}")
        super().__init__(f"{where}: {msg} (rc={rc})")
        self.rc = rc


class Desc(ctypes.Structure):
    _fields_ = [("nr", ctypes.c_int), ("ntheta", ctypes.c_int), ("nz", ctypes.c_int),
                ("halo", ctypes.c_int), ("lbound", ctypes.c_int), ("is_parallel", ctypes.c_int),
                ("periodic_theta", ctypes.c_int)]


class Counters(ctypes.Structure):
    """eafp_counters: process-wide activity counters of the library."""
    _fields_ = [("uploads", ctypes.c_longlong), ("labelled", ctypes.c_longlong),
                ("graph_captures", ctypes.c_longlong), ("graph_launches", ctypes.c_longlong)]


class Stats(ctypes.Structure):
    _fields_ = [("residual", ctypes.c_double), ("n_iter", ctypes.c_int), ("n_refine", ctypes.c_int),
                ("status", ctypes.c_int), ("pad_", ctypes.c_int), ("true_rel_l2", ctypes.c_double),
                ("true_rel_l1", ctypes.c_double), ("seconds", ctypes.c_double)]

    def as_dict(self):
        return {"residual": self.residual, "n_iter": self.n_iter, "n_refine": self.n_refine,
                "status": STOP_NAMES.get(self.status, str(self.status)),
                "true_rel_l2": self.true_rel_l2, "true_rel_l1": self.true_rel_l1,
                "seconds": self.seconds}


class MixedOpts(ctypes.Structure):
    _fields_ = [("tol", ctypes.c_double), ("interval", ctypes.c_int), ("max_steps", ctypes.c_int),
                ("coef_format", ctypes.c_int), ("restart", ctypes.c_int), ("precond", ctypes.c_int)]


STEP_FP64, STEP_MIXED = 0, 1


class PPStepOpts(ctypes.Structure):
    """eafp_pp_step_opts: the solver of the fused p' step and its options."""
    _fields_ = [("solver", ctypes.c_int), ("max_iter", ctypes.c_int), ("tol", ctypes.c_double),
                ("alpha_p", ctypes.c_double), ("mixed", MixedOpts)]


_lib = None
_dp = ctypes.POINTER(ctypes.c_double)


def load():
    """Load libeafp.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    L = ctypes.CDLL(LIB_PATH)
    i, d, vp = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
    L.eafp_cg_solve.argtypes = [ctypes.POINTER(Desc)] + [vp] * 9 + [i, d, _dp, ctypes.POINTER(i)]
    L.eafp_cg_solve.restype = i
    L.eafp_release.argtypes = []
    L.eafp_release.restype = i
    L.eafp_batch_create.argtypes = [i, ctypes.POINTER(Desc), ctypes.POINTER(vp)]
    L.eafp_batch_create.restype = i
    L.eafp_batch_upload.argtypes = [vp, i] + [vp] * 9 + [i]
    L.eafp_batch_upload.restype = i
    L.eafp_batch_solve_fp64.argtypes = [vp, i, d, ctypes.POINTER(Stats)]
    L.eafp_batch_solve_fp64.restype = i
    L.eafp_batch_solve_mixed.argtypes = [vp, ctypes.POINTER(MixedOpts), ctypes.POINTER(Stats)]
    L.eafp_batch_solve_mixed.restype = i
    L.eafp_batch_download.argtypes = [vp, i, vp, i]
    L.eafp_batch_download.restype = i
    L.eafp_batch_info.argtypes = [vp, ctypes.POINTER(i), ctypes.POINTER(i), ctypes.POINTER(ctypes.c_long)]
    L.eafp_batch_info.restype = i
    fp = ctypes.POINTER(ctypes.c_float)
    ip = ctypes.POINTER(i)
    L.eafp_batch_set_tiling.argtypes = [vp, i, i]
    L.eafp_batch_set_tiling.restype = i
    L.eafp_pp_assemble.argtypes = [ctypes.POINTER(Desc), ctypes.POINTER(PPFields), ctypes.POINTER(PPConfig)] + \
        [vp] * 8 + [_dp, _dp, ctypes.POINTER(PPExports), i]
    L.eafp_pp_assemble.restype = i
    L.eafp_pp_assemble_noalloc.argtypes = [ctypes.POINTER(Desc), ctypes.POINTER(PPFields), ctypes.POINTER(PPConfig)] + \
        [vp] * 8 + [_dp, _dp, ctypes.POINTER(PPExports), vp]
    L.eafp_pp_assemble_noalloc.restype = i
    L.eafp_pp_correct.argtypes = [ctypes.POINTER(Desc), ctypes.POINTER(PPFields), ctypes.POINTER(PPConfig), d, vp,
                                  ctypes.POINTER(PPExports), ctypes.POINTER(PPUpdate), i]
    L.eafp_pp_correct.restype = i
    L.eafp_pp_step.argtypes = [ctypes.POINTER(Desc), ctypes.POINTER(PPFields), ctypes.POINTER(PPConfig),
                               ctypes.POINTER(PPStepOpts), vp, ctypes.POINTER(PPUpdate), _dp, ctypes.POINTER(Stats), i]
    L.eafp_pp_step.restype = i
    L.eafp_tdma_solve.argtypes = [ctypes.POINTER(Desc)] + [vp] * 9 + [i, i]
    L.eafp_tdma_solve.restype = i
    L.eafp_tdma_solve_batch.argtypes = [i, ctypes.POINTER(Desc), ctypes.POINTER(TdmaSystem), i, i]
    L.eafp_tdma_solve_batch.restype = i
    L.eafp_batch_set_columns.argtypes = [vp, i]
    L.eafp_batch_set_columns.restype = i
    L.eafp_batch_get_columns.argtypes = [vp, ip]
    L.eafp_batch_get_columns.restype = i
    L.eafp_batch_time_kernels.argtypes = [vp, i, i, i, fp, fp]
    L.eafp_batch_time_kernels.restype = i
    L.eafp_batch_autotune.argtypes = [vp, i, i, i, ip, ip, fp]
    L.eafp_batch_autotune.restype = i
    L.eafp_batch_components.argtypes = [vp, i, ip, ip]
    L.eafp_batch_components.restype = i
    L.eafp_get_counters.argtypes = [ctypes.POINTER(Counters)]
    L.eafp_get_counters.restype = i
    L.eafp_batch_destroy.argtypes = [vp]
    L.eafp_batch_destroy.restype = i
    L.eafp_version.argtypes = []
    L.eafp_version.restype = ctypes.c_char_p
    _lib = L
    return L


def check(rc, where):
    if rc != EAFP_OK:
        raise EafpError(rc, where)
