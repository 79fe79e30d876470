"""Pins the oracle's restatement of the p' assembly (oracle/pp_assemble.c,
src/mod_pressure_3d.f90:171-306 and src/mod_boundary_3d.f90:180-277) with
properties the mathematics fixes, not with its own formulas:

* every face coefficient is written identically from both sides: a_E(i) =
  a_W(i+1), a_N(j) = a_S(j+1), a_T(k) = a_B(k+1), exactly, and across the
  periodic seam to the last bits (a wrong area, distance or neighbour index on
  one face breaks it);
* the face fluxes telescope: on a domain without outlets or compressibility
  source, sum(Su) = 0 to rounding (a wrong sign on one face breaks it);
* the rest state: zero velocities with a pressure that is linear in z and
  uniform phases gives u_f = 0 on every face, so Su = 0 (the Rhie-Chow term
  and the liquid-weight correction of the gas faces cancel exactly when
  their signs and averages are right);
* the diagonal: every active row has aP >= (1 + 1e-6) * sum(a_nb), identity
  rows for masked cells, cells without an active phase and outlet cells;
* the boundary rows: Neumann walls at i = 1, nr and k = 1, nz, the Dirichlet
  folding around outlets (no link into an outlet cell).
"""
import numpy as np
import pytest

import oracle
from eafp import fields

NAMES = ("aW", "aE", "aS", "aN", "aB", "aT", "aP", "Su")


def interior(a):
    return a[2:-2, 2:-2, 2:-2]


def cases():
    return [((12, 24, 16), dict(seed=0)), ((20, 36, 28), dict(seed=1, masks=True)),
            ((16, 30, 24), dict(seed=2, two_phase=False)), ((24, 48, 32), dict(seed=3, scrap=False))]


@pytest.mark.parametrize("shape,kw", cases())
def test_face_coefficients_symmetric(shape, kw):
    F, cfg = fields.make_fields(*shape, **kw)
    out, res, flux = oracle.pp_assemble(F, cfg, shape)
    aW, aE, aS, aN, aB, aT = (interior(out[n]) for n in NAMES[:6])
    assert np.count_nonzero(aE) > 0.05 * aE.size
    np.testing.assert_array_equal(aE[:-1], aW[1:])
    np.testing.assert_array_equal(aN[:, :-1], aS[:, 1:])
    # across the periodic seam the two sides difference the unrolled theta
    # coordinates at opposite ends (theta(nth+1) - theta(nth) against
    # theta(1) - theta(0)): equal to the last bits only, as in the reference
    np.testing.assert_allclose(aN[:, -1], aS[:, 0], rtol=1e-13, atol=0)
    np.testing.assert_array_equal(aT[:, :, :-1], aB[:, :, 1:])
    assert np.isfinite(res) and flux > 0


@pytest.mark.parametrize("shape,kw", cases())
def test_face_fluxes_telescope(shape, kw):
    F, cfg = fields.make_fields(*shape, **kw)
    F["outlet"] = np.zeros_like(F["outlet"])
    cfg = dict(cfg, gas_compressibility=0)
    out, res, flux = oracle.pp_assemble(F, cfg, shape)
    su = interior(out["Su"])
    assert np.abs(su).sum() > 1e-3 * flux           # a real imbalance ...
    assert abs(su.sum()) <= 1e-12 * flux            # ... that sums to zero over the closed domain


@pytest.mark.parametrize("two_phase", [False, True])
def test_rest_state_has_no_source(two_phase):
    shape = (14, 24, 20)
    F, cfg = fields.make_fields(*shape, seed=4, two_phase=two_phase, scrap=False)
    fluid = F["cell_type"] != 0
    for n in ("ur_l", "uth_l", "uz_l", "ur_g", "uth_g", "uz_g"):
        F[n] = np.zeros_like(F[n], order="F")
    # uniform phases and properties, pressure linear in z
    F["alpha_l"] = np.asfortranarray(np.where(fluid, 0.6, 0.0))
    F["alpha_g"] = np.asfortranarray(np.where(fluid, 0.4, 0.0))
    for n, v in (("rho_l", 7000.0), ("T_l", 1850.0), ("rho_g", 0.3)):
        F[n] = np.full_like(F[n], v, order="F")
    z = F["z"]
    F["p"] = np.asfortranarray(np.broadcast_to(-3000.0 * z[None, None, :], F["p"].shape).copy())
    F["outlet"] = np.zeros_like(F["outlet"])
    cfg = dict(cfg, gas_compressibility=0)
    out, res, flux = oracle.pp_assemble(F, cfg, shape)
    su = interior(out["Su"])
    # u_f = d_f (g_avg - dp/dz) vanishes to rounding; compare with the size of one face term
    scale = np.abs(interior(out["aT"])).max() * 3000.0 * 0.3
    assert np.abs(su).max() <= 1e-10 * scale


@pytest.mark.parametrize("shape,kw", cases())
def test_diagonal_and_boundary_rows(shape, kw):
    F, cfg = fields.make_fields(*shape, **kw)
    out, res, flux = oracle.pp_assemble(F, cfg, shape)
    a = {n: interior(out[n]) for n in NAMES}
    nbsum = a["aW"] + a["aE"] + a["aS"] + a["aN"] + a["aB"] + a["aT"]
    ct = interior(F["cell_type"]) != 0
    act_l = interior(F["liq_cont"]) != 0 if "liq_cont" in F else interior(F["alpha_l"]) >= 1e-2
    act_g = interior(F["alpha_g"]) >= 1e-2
    if "pv_active" in F:
        act_g &= interior(F["pv_active"]) != 0
    if "gas_reach" in F:
        act_g |= (interior(F["gas_reach"]) != 0) & (interior(F["alpha_g"]) < 1e-2) & ct
    # the trivial-row test uses both activity masks whatever the phase flags (:249-259)
    outlet_top = F["outlet"][2:-2, 2:-2] != 0
    identity = ~ct | ~(act_l | act_g)
    identity[:, :, -1] |= outlet_top
    for n in NAMES[:6]:
        assert not a[n][identity].any(), n
    assert (a["aP"][identity] == 1.0).all() and not a["Su"][identity].any()
    rows = ~identity
    assert (a["aP"][rows] >= nbsum[rows] * (1.0 + 1e-6)).all()
    assert not a["aW"][0].any() and not a["aE"][-1].any()
    assert not a["aB"][:, :, 0].any() and not a["aT"][:, :, -1].any()
    # Dirichlet folding: no link into an outlet cell
    o = outlet_top
    assert not a["aT"][:, :, -2][o].any()
    assert not a["aE"][:-1, :, -1][o[1:]].any() and not a["aW"][1:, :, -1][o[:-1]].any()
    assert not a["aN"][:, :, -1][np.roll(o, -1, axis=1)].any()
    assert not a["aS"][:, :, -1][np.roll(o, 1, axis=1)].any()


def test_single_phase_is_symmetric_positive_definite():
    """The assembled operand (walls, outlet anchor, compliance) is a valid CG
    input: the oracle CG converges on it."""
    shape = (12, 24, 16)
    F, cfg = fields.make_fields(*shape, seed=5)
    out, res, flux = oracle.pp_assemble(F, cfg, shape)

    class S:
        pass
    ps = S()
    ps.shape = shape
    ps.padded = lambda n: out[n]
    phi, r, nit, _ = oracle.cg(ps, max_iter=20000, tol=1e-10)
    assert r < 1e-10 and 0 < nit < 20000
