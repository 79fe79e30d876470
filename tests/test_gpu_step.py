"""GPU parity of the fused p' step (eafp_pp_step): assembly, solve and
post-solve steps in one call with the system kept in HBM.

Bars:
  - FP64 solver: every output (p', fluxes, link flags, velocities, p, the
    continuity residual, the solver's residual and iteration count) is
    bit-identical to the three entry points called in turn
    (eafp_pp_assemble, eafp_cg_solve from p' = 0, eafp_pp_correct), each of
    which is held to the oracle by its own tests; and, converged, the
    outputs agree with the oracle's three steps to the solution's accuracy
    (relative change error <= 1e-6, as for the unfused pipeline).
  - Mixed solver: the returned p' leaves a true relative residual below the
    requested 1e-8 on the oracle-assembled system (computed here in FP64
    numpy), and the post-solve outputs equal eafp_pp_correct applied to
    that p' bitwise.
  - Host and device arrays, and in/out arrays that alias the fields, give
    the same bits.
"""
import ctypes

import numpy as np
import pytest

import eafp
import oracle
from eafp import _lib, fields
from hipbuf import HipBuffers

pytestmark = pytest.mark.gpu

NAMES = ("aW", "aE", "aS", "aN", "aB", "aT", "aP", "Su")
UPD = _lib.PP_UPDATE


def _three_calls(F, cfg, m, alpha_p, max_iter, tol):
    out, res, _, ex = eafp.pp_assemble(F, cfg, m, exports=True)
    pp = np.zeros_like(out["aP"], order="F")
    r_cg, n_cg = eafp.cg_3d_mpi(*[out[n] for n in NAMES], pp, m, max_iter, tol)
    if r_cg != r_cg:
        res = r_cg
    upd = eafp.pp_correct(F, cfg, m, alpha_p, pp, ex)
    return upd, pp, res, r_cg, n_cg


def _assert_same(a, b, what):
    for n in UPD:
        np.testing.assert_array_equal(a[n], b[n], err_msg=f"{what}: {n}")


STEP_CASES = [((12, 24, 16), dict(seed=0), 0.3, 50),           # the reference's max_inner_pres
              ((20, 36, 28), dict(seed=1, masks=True), 0.5, 50),
              ((16, 30, 24), dict(seed=2, two_phase=False), 0.3, 400),
              ((9, 2, 7), dict(seed=3, masks=True), 0.7, 50),
              ((60, 120, 84), dict(seed=5, masks=True), 0.3, 50)]


@pytest.mark.parametrize("shape,kw,alpha_p,max_iter", STEP_CASES)
def test_step_fp64_equals_three_calls(shape, kw, alpha_p, max_iter):
    print(f"case step-fp64 shape={shape} {kw} alpha_p={alpha_p} max_iter={max_iter}", flush=True)
    F, cfg = fields.make_fields(*shape, **kw)
    m = eafp.Mesh(*shape)
    ref, pp_ref, res_ref, r_cg, n_cg = _three_calls(F, cfg, m, alpha_p, max_iter, 1e-8)
    out, pp, res, st = eafp.pp_step(F, cfg, m, alpha_p, max_iter=max_iter, tol=1e-8)
    _assert_same(out, ref, "step vs three calls")
    np.testing.assert_array_equal(pp, pp_ref)
    assert res == res_ref
    assert st["residual"] == r_cg and st["n_iter"] == n_cg
    assert np.count_nonzero(pp) > 0 and np.count_nonzero(out["p"] != F["p"]) > 0
    print(f"  n_iter {st['n_iter']} cg residual {st['residual']:.3e} continuity residual {res:.6e}", flush=True)


def test_step_fp64_converged_matches_oracle_steps():
    """Converged, the fused step agrees with the oracle's assemble -> CG -> correct."""
    shape = (20, 36, 28)
    F, cfg = fields.make_fields(*shape, seed=9, masks=True)
    out, pp, res, st = eafp.pp_step(F, cfg, eafp.Mesh(*shape), 0.3, max_iter=20000, tol=1e-11)
    ref, res_ref, _, ex_ref = oracle.pp_assemble(F, cfg, shape, exports=True)

    class S:
        pass
    ps = S()
    ps.shape = shape
    ps.padded = lambda n: ref[n]
    pp_ref, r_ref, _, _ = oracle.cg(ps, max_iter=20000, tol=1e-11)
    upd_ref = oracle.pp_correct(F, cfg, shape, 0.3, pp_ref, ex_ref)
    print(f"case step-converged: gpu res={st['residual']:.2e} n={st['n_iter']} oracle res={r_ref:.2e}", flush=True)
    assert st["status"] == "converged" and st["residual"] < 1e-11
    assert abs(res - res_ref) <= 1e-12 * max(res_ref, 1e-300)
    x, xr = oracle.interior(pp), oracle.interior(pp_ref)
    assert np.linalg.norm(x - xr) <= 1e-6 * np.linalg.norm(xr)
    for n in ("lk_r", "lk_th", "lk_z"):
        np.testing.assert_array_equal(out[n], upd_ref[n], err_msg=n)
    for n in ("Fc_r", "Fc_th", "Fc_z", "ur_l", "uth_l", "uz_l", "ur_g", "uth_g", "uz_g", "p"):
        err = np.linalg.norm(out[n] - upd_ref[n]) / max(np.linalg.norm(upd_ref[n] - F.get(n, 0.0)), 1e-300)
        assert err <= 1e-6, f"{n}: relative change error {err:.2e}"


def _true_rel_residual(A, pp, shape):
    """||Su - A pp|| / ||Su|| over the interior, the operator applied in FP64
    numpy from the padded coefficient arrays (theta neighbours through the
    periodic halos of pp, r/z neighbours through its zero halos)."""
    nr, nth, nz = shape
    I = (slice(2, nr + 2), slice(2, nth + 2), slice(2, nz + 2))

    def sh(di, dj, dk):
        return pp[2 + di:nr + 2 + di, 2 + dj:nth + 2 + dj, 2 + dk:nz + 2 + dk]
    Ax = (A["aP"][I] * sh(0, 0, 0) - A["aW"][I] * sh(-1, 0, 0) - A["aE"][I] * sh(1, 0, 0)
          - A["aS"][I] * sh(0, -1, 0) - A["aN"][I] * sh(0, 1, 0) - A["aB"][I] * sh(0, 0, -1)
          - A["aT"][I] * sh(0, 0, 1))
    b = A["Su"][I]
    return np.linalg.norm(b - Ax) / np.linalg.norm(b)


@pytest.mark.parametrize("shape,kw,fmt", [((20, 36, 28), dict(seed=1, masks=True), _lib.COEF_BF16),
                                          ((16, 30, 24), dict(seed=2, two_phase=False), _lib.COEF_FP16),
                                          ((60, 120, 84), dict(seed=5, masks=True), _lib.COEF_BF16)])
def test_step_mixed(shape, kw, fmt):
    print(f"case step-mixed shape={shape} {kw} fmt={fmt}", flush=True)
    F, cfg = fields.make_fields(*shape, **kw)
    m = eafp.Mesh(*shape)
    out, pp, res, st = eafp.pp_step(F, cfg, m, 0.3, solver="mixed", mixed=dict(tol=1e-8, coef_format=fmt))
    ref, res_ref, _, ex_ref = oracle.pp_assemble(F, cfg, shape, exports=True)
    rel = _true_rel_residual(ref, pp, shape)
    print(f"  status {st['status']} inner {st['n_iter']} refine {st['n_refine']} "
          f"reported {st['residual']:.2e} numpy {rel:.2e}", flush=True)
    assert st["status"] == "converged"
    assert rel < 1e-8 * 1.01
    assert abs(res - res_ref) <= 1e-12 * max(res_ref, 1e-300)
    # theta halos filled, r/z halos zero
    nr, nth, nz = shape
    np.testing.assert_array_equal(pp[:, 0, :], pp[:, nth, :])
    np.testing.assert_array_equal(pp[:, nth + 2, :], pp[:, 2, :])
    assert not pp[:2].any() and not pp[nr + 2:].any() and not pp[:, :, :2].any() and not pp[:, :, nz + 2:].any()
    _, _, _, ex = eafp.pp_assemble(F, cfg, m, exports=True)
    _assert_same(out, eafp.pp_correct(F, cfg, m, 0.3, pp, ex), "mixed step vs pp_correct of its p'")


def test_step_in_place_aliasing():
    """liq%ur and sh%p are both inputs of the assembly and outputs of the
    corrections: passing the same arrays gives the same bits."""
    shape = (14, 24, 20)
    F, cfg = fields.make_fields(*shape, seed=4, masks=True)
    m = eafp.Mesh(*shape)
    out, pp, res, st = eafp.pp_step(F, cfg, m, 0.4, max_iter=50)
    G = {n: (np.array(v, order="F") if isinstance(v, np.ndarray) else v) for n, v in F.items()}
    for n in UPD[6:]:
        G[n] = np.asfortranarray(G[n], dtype=np.float64)
    out2, pp2, res2, st2 = eafp.pp_step(G, cfg, m, 0.4, max_iter=50, in_place=True)
    _assert_same(out2, out, "in place vs copies")
    for n in UPD[6:]:
        assert out2[n] is G[n]
    np.testing.assert_array_equal(pp2, pp)
    assert res2 == res and st2["n_iter"] == st["n_iter"]


def test_step_device_resident():
    """arrays_on_device = 1: fields, io and pp all device pointers."""
    shape = (20, 36, 28)
    F, cfg = fields.make_fields(*shape, seed=8, masks=True)
    m = eafp.Mesh(*shape)
    ref, pp_ref, res_ref, st_ref = eafp.pp_step(F, cfg, m, 0.3, max_iter=50)
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
        upd = {n: hip.zeros((4 if isint(n) else 8) * npad) for n in UPD[:6]}
        # the velocities and p alias the fields' device arrays, as in a GPU-resident caller
        upd.update({n: vals[n] for n in UPD[6:]})
        io = _lib.PPUpdate(**upd)
        pp_d = hip.zeros(8 * npad)
        o = _lib.PPStepOpts(_lib.STEP_FP64, 50, 1e-8, 0.3, _lib.MixedOpts(1e-8, 32, 2000, _lib.COEF_BF16, 0))
        res = ctypes.c_double()
        st = _lib.Stats()
        L = _lib.load()
        d = eafp.solver.desc_of(m)
        _lib.check(L.eafp_pp_step(ctypes.byref(d), ctypes.byref(f), ctypes.byref(c), ctypes.byref(o), pp_d,
                                  ctypes.byref(io), ctypes.byref(res), ctypes.byref(st), 1), "eafp_pp_step")
        for n in UPD:
            np.testing.assert_array_equal(hip.get(upd[n], pad, np.int32 if isint(n) else np.float64), ref[n],
                                          err_msg=n)
        np.testing.assert_array_equal(hip.get(pp_d, pad, np.float64), pp_ref)
        assert res.value == res_ref and st.n_iter == st_ref["n_iter"]
    finally:
        hip.free()


def test_step_nan_residual_propagates():
    """A NaN in the fields: the step still returns OK, reports a NaN
    residual (src/mod_pressure_3d.f90:322) and applies the corrections, bit
    for bit as the three calls do."""
    shape = (12, 24, 16)
    F, cfg = fields.make_fields(*shape, seed=0)
    F = dict(F)
    F["vol"] = np.array(F["vol"], order="F")
    nr, nth, nz = shape
    ijk = np.argwhere(F["cell_type"][2:nr + 2, 2:nth + 2, 2:nz + 2] != 0)[100] + 2
    F["vol"][tuple(ijk)] = np.nan
    m = eafp.Mesh(*shape)
    ref, pp_ref, res_ref, r_cg, _ = _three_calls(F, cfg, m, 0.3, 50, 1e-8)
    out, pp, res, st = eafp.pp_step(F, cfg, m, 0.3, max_iter=50)
    assert np.isnan(res) and np.isnan(st["residual"]) and np.isnan(r_cg) and np.isnan(res_ref)
    _assert_same(out, ref, "NaN step vs three calls")
    np.testing.assert_array_equal(pp, pp_ref)
