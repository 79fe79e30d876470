"""GPU parity of the p' assembler (eafp_pp_assemble) against the oracle's
restatement (oracle/pp_assemble.c, pinned by tests/test_pp_oracle.py).

Bar: every coefficient array and Su bit-identical (integer-like work on
doubles: the same operations in the same order, no contraction); the
continuity residual and flux_ref, whose global sums are reduced in another
order, within 1e-12 relative.  End to end, the GPU-assembled system solved
by the GPU drop-in CG matches the oracle CG on the oracle-assembled system.
"""
import numpy as np
import pytest

import eafp
import oracle
from eafp import fields
from hipbuf import HipBuffers

pytestmark = pytest.mark.gpu

NAMES = ("aW", "aE", "aS", "aN", "aB", "aT", "aP", "Su")
CASES = [((12, 24, 16), dict(seed=0)),
         ((20, 36, 28), dict(seed=1, masks=True)),
         ((16, 30, 24), dict(seed=2, two_phase=False)),
         ((24, 48, 32), dict(seed=3, scrap=False)),
         ((7, 12, 9), dict(seed=4, masks=True)),
         ((60, 120, 84), dict(seed=5, masks=True))]


@pytest.mark.parametrize("shape,kw", CASES)
def test_assembly_matches_oracle_bitwise(shape, kw):
    print(f"case assemble shape={shape} {kw}", flush=True)
    F, cfg = fields.make_fields(*shape, **kw)
    ref, res_ref, flux_ref = oracle.pp_assemble(F, cfg, shape)
    out, res, flux = eafp.pp_assemble(F, cfg, eafp.Mesh(*shape))
    for n in NAMES:
        bad = np.count_nonzero(out[n] != ref[n])
        assert bad == 0, f"{n}: {bad} cells differ, max |diff| {np.abs(out[n] - ref[n]).max():.3e}"
    assert abs(flux - flux_ref) <= 1e-12 * flux_ref
    assert abs(res - res_ref) <= 1e-12 * max(res_ref, 1e-300)
    print(f"  residual {res:.6e} (oracle {res_ref:.6e}) flux_ref {flux:.6e}", flush=True)


@pytest.mark.parametrize("flags", [dict(gas_in_poisson=0), dict(gas_compressibility=0), dict(solve_multiphase=0)])
def test_assembly_flags(flags):
    shape = (14, 24, 20)
    F, cfg = fields.make_fields(*shape, seed=6, masks=True)
    cfg = dict(cfg, **flags)
    ref, res_ref, _ = oracle.pp_assemble(F, cfg, shape)
    out, res, _ = eafp.pp_assemble(F, cfg, eafp.Mesh(*shape))
    for n in NAMES:
        np.testing.assert_array_equal(out[n], ref[n], err_msg=n)
    assert abs(res - res_ref) <= 1e-12 * max(res_ref, 1e-300)


def test_assemble_then_solve_matches_oracle_pipeline():
    """GPU assembly -> GPU drop-in CG against oracle assembly -> oracle CG."""
    shape = (20, 36, 28)
    F, cfg = fields.make_fields(*shape, seed=7, masks=True)
    m = eafp.Mesh(*shape)
    out, res, _ = eafp.pp_assemble(F, cfg, m)
    phi = np.zeros_like(out["aP"], order="F")
    r_gpu, n_gpu = eafp.cg_3d_mpi(*[out[n] for n in NAMES], phi, m, 20000, 1e-10)
    ref, _, _ = oracle.pp_assemble(F, cfg, shape)

    class S:
        pass
    ps = S()
    ps.shape = shape
    ps.padded = lambda n: ref[n]
    phi_ref, r_ref, n_ref, _ = oracle.cg(ps, max_iter=20000, tol=1e-10)
    print(f"case pipeline: gpu n_iter={n_gpu} res={r_gpu:.2e} | oracle n_iter={n_ref} res={r_ref:.2e}", flush=True)
    assert r_gpu < 1e-10 and abs(n_gpu - n_ref) <= max(2, 0.02 * n_ref)
    x, xr = oracle.interior(phi), oracle.interior(phi_ref)
    assert np.linalg.norm(x - xr) <= 1e-6 * np.linalg.norm(xr)


def _smooth_pp(shape, seed, scale=50.0):
    """A seeded correction field with periodic theta halos, as the CG leaves it."""
    nr, nth, nz = shape
    rng = np.random.default_rng(900 + seed)
    pp = np.asfortranarray(scale * rng.normal(size=(nr + 4, nth + 4, nz + 4)))
    return fields.theta_halo(pp)


def _sealed_cells(F, shape):
    ct, al, ag = F["cell_type"], F["alpha_l"], F["alpha_g"]
    act_l = F["liq_cont"] != 0 if "liq_cont" in F else al >= 1e-2
    act_g = ag >= 1e-2
    if "pv_active" in F:
        act_g &= F["pv_active"] != 0
    if "gas_reach" in F:
        act_g |= (F["gas_reach"] != 0) & (ag < 1e-2) & (ct != 0)
    s = (ct != 0) & ~act_l & ~act_g
    nr, nth, nz = shape
    return int(np.count_nonzero(s[2:nr + 2, 2:nth + 2, 2:nz + 2]))


UPD = ("Fc_r", "Fc_th", "Fc_z", "lk_r", "lk_th", "lk_z", "ur_l", "uth_l", "uz_l", "ur_g", "uth_g", "uz_g", "p")
CORRECT_CASES = [((12, 24, 16), dict(seed=0), 0.3, 50.0),
                 ((20, 36, 28), dict(seed=1, masks=True), 0.5, 50.0),
                 ((16, 30, 24), dict(seed=2, two_phase=False), 0.3, 50.0),
                 ((9, 2, 7), dict(seed=3, masks=True), 0.7, 50.0),
                 ((14, 24, 20), dict(seed=4, masks=True), 1.0, 3.0e6),   # the +-2e6 clamp engages
                 ((60, 120, 84), dict(seed=5, masks=True), 0.3, 50.0)]


@pytest.mark.parametrize("shape,kw", CASES[:4])
def test_exports_match_oracle_bitwise(shape, kw):
    print(f"case exports shape={shape} {kw}", flush=True)
    F, cfg = fields.make_fields(*shape, **kw)
    ref, _, _, ex_ref = oracle.pp_assemble(F, cfg, shape, exports=True)
    out, _, _, ex = eafp.pp_assemble(F, cfg, eafp.Mesh(*shape), exports=True)
    for n in NAMES:
        np.testing.assert_array_equal(out[n], ref[n], err_msg=n)
    for n in ex_ref:
        bad = np.count_nonzero(ex[n] != ex_ref[n])
        assert bad == 0, f"{n}: {bad} cells differ"
    assert np.count_nonzero(ex["lk_z"]) > 0


@pytest.mark.parametrize("shape,kw,alpha_p,scale", CORRECT_CASES)
def test_post_solve_matches_oracle_bitwise(shape, kw, alpha_p, scale):
    print(f"case correct shape={shape} {kw} alpha_p={alpha_p} scale={scale}", flush=True)
    F, cfg = fields.make_fields(*shape, **kw)
    _, _, _, ex = oracle.pp_assemble(F, cfg, shape, exports=True)
    pp = _smooth_pp(shape, kw["seed"], scale)
    ref = oracle.pp_correct(F, cfg, shape, alpha_p, pp, ex)
    out = eafp.pp_correct(F, cfg, eafp.Mesh(*shape), alpha_p, pp, ex)
    for n in UPD:
        bad = np.count_nonzero(out[n] != ref[n])
        assert bad == 0, f"{n}: {bad} cells differ"
    # the case exercises what it claims: velocities and p changed, sealed cells present
    assert np.count_nonzero(out["ur_l"] != F["ur_l"]) > 0
    assert np.count_nonzero(out["p"] != F["p"]) > 0
    if kw.get("two_phase", True):
        assert np.count_nonzero(out["uz_g"] != F["uz_g"]) > 0
    else:
        np.testing.assert_array_equal(out["uz_g"], F["uz_g"])
    if scale > 1e6:
        assert np.count_nonzero(np.abs(out["p"]) == 2.0e6) > 0
    sealed = _sealed_cells(F, shape)
    print(f"  sealed cells {sealed}, linked z faces {int(np.count_nonzero(out['lk_z']))}", flush=True)
    if kw.get("masks"):
        assert sealed > 0


def test_device_resident_assemble_and_correct():
    """arrays_on_device = 1: every array pointer a device pointer."""
    import ctypes
    from eafp import _lib
    shape = (20, 36, 28)
    F, cfg = fields.make_fields(*shape, seed=8, masks=True)
    ref, res_ref, _, ex_ref = oracle.pp_assemble(F, cfg, shape, exports=True)
    pp = _smooth_pp(shape, 8)
    upd_ref = oracle.pp_correct(F, cfg, shape, 0.4, pp, ex_ref)
    pad = tuple(n + 4 for n in shape)
    npad = int(np.prod(pad))
    hip = HipBuffers()
    try:
        vals = {n: (hip.put(np.asarray(F[n], dtype=np.int32 if n in _lib.PP_INT else np.float64)) if n in F else None)
                for n in _lib.PP_FIELDS}
        f = _lib.PPFields(**vals)
        c = _lib.PPConfig(cfg["dt"], cfg["rho_gas"], cfg["T_ambient"], cfg["beta_expansion"],
                          cfg["gas_in_poisson"], cfg["solve_multiphase"], cfg["gas_compressibility"], 0)
        isint = lambda n: n.startswith("lk")
        outs = {n: hip.zeros(8 * npad) for n in NAMES}
        ex = {n: hip.zeros((4 if isint(n) else 8) * npad) for n in _lib.PP_EXPORTS}
        exs = _lib.PPExports(**ex)
        L = _lib.load()
        d = eafp.solver.desc_of(eafp.Mesh(*shape))
        res = ctypes.c_double()
        _lib.check(L.eafp_pp_assemble(ctypes.byref(d), ctypes.byref(f), ctypes.byref(c), *[outs[n] for n in NAMES],
                                      ctypes.byref(res), None, ctypes.byref(exs), 1), "eafp_pp_assemble")
        for n in NAMES:
            np.testing.assert_array_equal(hip.get(outs[n], pad, np.float64), ref[n], err_msg=n)
        for n in _lib.PP_EXPORTS:
            np.testing.assert_array_equal(hip.get(ex[n], pad, np.int32 if isint(n) else np.float64), ex_ref[n],
                                          err_msg=n)
        assert abs(res.value - res_ref) <= 1e-12 * max(res_ref, 1e-300)
        upd = {n: hip.zeros((4 if isint(n) else 8) * npad) for n in _lib.PP_UPDATE[:6]}
        upd.update({n: hip.put(np.asarray(F[n], dtype=np.float64)) for n in _lib.PP_UPDATE[6:]})
        io = _lib.PPUpdate(**upd)
        pp_d = hip.put(pp)
        _lib.check(L.eafp_pp_correct(ctypes.byref(d), ctypes.byref(f), ctypes.byref(c), 0.4, pp_d,
                                     ctypes.byref(exs), ctypes.byref(io), 1), "eafp_pp_correct")
        for n in _lib.PP_UPDATE:
            np.testing.assert_array_equal(hip.get(upd[n], pad, np.int32 if isint(n) else np.float64), upd_ref[n],
                                          err_msg=n)
    finally:
        hip.free()


def test_assemble_solve_correct_pipeline():
    """GPU assembly -> GPU CG -> GPU post-solve against the oracle's three
    steps: the p' systems agree bitwise, the two CGs to their tolerance, so
    the consumers agree to the solution's accuracy."""
    shape = (20, 36, 28)
    F, cfg = fields.make_fields(*shape, seed=9, masks=True)
    m = eafp.Mesh(*shape)
    out, _, _, ex = eafp.pp_assemble(F, cfg, m, exports=True)
    pp = np.zeros_like(out["aP"], order="F")
    r_gpu, _ = eafp.cg_3d_mpi(*[out[n] for n in NAMES], pp, m, 20000, 1e-11)
    upd = eafp.pp_correct(F, cfg, m, 0.3, pp, ex)
    ref, _, _, ex_ref = oracle.pp_assemble(F, cfg, shape, exports=True)

    class S:
        pass
    ps = S()
    ps.shape = shape
    ps.padded = lambda n: ref[n]
    pp_ref, r_ref, _, _ = oracle.cg(ps, max_iter=20000, tol=1e-11)
    upd_ref = oracle.pp_correct(F, cfg, shape, 0.3, pp_ref, ex_ref)
    print(f"case pipeline-correct: gpu res={r_gpu:.2e} oracle res={r_ref:.2e}", flush=True)
    assert r_gpu < 1e-11
    for n in ("lk_r", "lk_th", "lk_z"):
        np.testing.assert_array_equal(upd[n], upd_ref[n], err_msg=n)
    for n in ("Fc_r", "Fc_th", "Fc_z", "ur_l", "uth_l", "uz_l", "ur_g", "uth_g", "uz_g", "p"):
        err = np.linalg.norm(upd[n] - upd_ref[n]) / max(np.linalg.norm(upd_ref[n] - F.get(n, 0.0)), 1e-300)
        assert err <= 1e-6, f"{n}: relative change error {err:.2e}"
