"""The allocation-free assembly (eafp_pp_assemble_noalloc) through the C-ABI
against the existing entry point (eafp_pp_assemble with arrays_on_device =
1) on the same device arrays: every coefficient, Su and face export bit for
bit, the continuity residual and flux_ref bit for bit (the block sums are
reduced on the device in the host's order).  Two calls in a row, the second
on a smaller mesh, so the shared scratch array holds the first call's
values when the second starts; the second call runs on a caller's stream.
A mesh beyond the scratch is refused with EAFP_ERR_SIZE before any device
work."""
import ctypes

import numpy as np
import pytest

import eafp
from eafp import _lib, fields
from hipbuf import HipBuffers

pytestmark = pytest.mark.gpu

NAMES = ("aW", "aE", "aS", "aN", "aB", "aT", "aP", "Su")


def _device_case(hip, shape, seed):
    F, cfg = fields.make_fields(*shape, seed=seed, masks=True)
    vals = {n: (hip.put(np.asarray(F[n], dtype=np.int32 if n in _lib.PP_INT else np.float64)) if n in F else None)
            for n in _lib.PP_FIELDS}
    f = _lib.PPFields(**vals)
    c = _lib.PPConfig(cfg["dt"], cfg["rho_gas"], cfg["T_ambient"], cfg["beta_expansion"],
                      cfg["gas_in_poisson"], cfg["solve_multiphase"], cfg["gas_compressibility"], 0)
    return f, c


def _outputs(hip, npad):
    isint = lambda n: n.startswith("lk")
    outs = {n: hip.zeros(8 * npad) for n in NAMES}
    ex = {n: hip.zeros((4 if isint(n) else 8) * npad) for n in _lib.PP_EXPORTS}
    return outs, ex


def _fill_nan(hip, ptr, nbytes):
    assert hip.h.hipMemset(ctypes.c_void_p(ptr), 0xFF, ctypes.c_size_t(nbytes)) == 0
    assert hip.h.hipDeviceSynchronize() == 0


def test_noalloc_assembly_matches_the_allocating_entry_point():
    L = _lib.load()
    hip = HipBuffers()
    stream = ctypes.c_void_p()
    assert hip.h.hipStreamCreate(ctypes.byref(stream)) == 0
    try:
        for call, (shape, seed, strm) in enumerate((((40, 72, 56), 21, None), ((20, 36, 28), 22, stream.value))):
            print(f"case noalloc call={call} shape={shape} seed={seed} stream={'caller' if strm else 'default'}",
                  flush=True)
            pad = tuple(n + 4 for n in shape)
            npad = int(np.prod(pad))
            assert npad <= _lib.NOALLOC_MAX_CELLS
            f, c = _device_case(hip, shape, seed)
            d = eafp.solver.desc_of(eafp.Mesh(*shape))
            outs, ex = _outputs(hip, npad)
            res, flux = ctypes.c_double(), ctypes.c_double()
            _lib.check(L.eafp_pp_assemble(ctypes.byref(d), ctypes.byref(f), ctypes.byref(c),
                                          *[outs[n] for n in NAMES], ctypes.byref(res), ctypes.byref(flux),
                                          ctypes.byref(_lib.PPExports(**ex)), 1), "eafp_pp_assemble")
            outs2, ex2 = _outputs(hip, npad)
            for n, p in list(outs2.items()) + list(ex2.items()):
                _fill_nan(hip, p, (4 if n.startswith("lk") else 8) * npad)
            res2, flux2 = ctypes.c_double(-1.0), ctypes.c_double(-1.0)
            _lib.check(L.eafp_pp_assemble_noalloc(ctypes.byref(d), ctypes.byref(f), ctypes.byref(c),
                                                  *[outs2[n] for n in NAMES], ctypes.byref(res2),
                                                  ctypes.byref(flux2), ctypes.byref(_lib.PPExports(**ex2)), strm),
                       "eafp_pp_assemble_noalloc")
            for n in NAMES:
                a = hip.get(outs[n], pad, np.float64)
                b = hip.get(outs2[n], pad, np.float64)
                assert a.tobytes(order="F") == b.tobytes(order="F"), n
            for n in _lib.PP_EXPORTS:
                dt = np.int32 if n.startswith("lk") else np.float64
                a = hip.get(ex[n], pad, dt)
                b = hip.get(ex2[n], pad, dt)
                assert a.tobytes(order="F") == b.tobytes(order="F"), n
            print(f"  residual {res2.value!r} (allocating {res.value!r}) flux_ref {flux2.value!r}", flush=True)
            assert res2.value == res.value and flux2.value == flux.value and flux.value > 0
            # beyond the scratch: refused before any device work
            big = eafp.solver.desc_of(eafp.Mesh(100, 100, 100))
            rc = L.eafp_pp_assemble_noalloc(ctypes.byref(big), ctypes.byref(f), ctypes.byref(c),
                                            *[outs2[n] for n in NAMES], None, None, None, None)
            assert rc == _lib.EAFP_ERR_SIZE
    finally:
        hip.h.hipStreamSynchronize(stream)
        hip.h.hipStreamDestroy(stream)
        hip.free()
