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

and its post-solve consumers (oracle_pp_correct, src/mod_pressure_3d.f90:324-397,
:689-746):

* the conservative fluxes of the solved system balance: each cell's net
  outflow equals -(aP - sum a_nb) pp_P (a wrong sign, neighbour or export
  breaks it);
* a p' linear along one axis moves only that axis's velocities, by the
  closed form -alpha V / aP_u * slope (walls excepted in r and z);
* the pressure update is the clamped p + alpha_p pp outside sealed cells,
  each sealed cell lies within the range of its active neighbours (0 without
  one), and the theta halos are periodic copies.
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


# ---------------------------------------------------------------------------
# Post-solve consumers (oracle_pp_correct: src/mod_pressure_3d.f90:333-397,
# 689-746)
# ---------------------------------------------------------------------------

def _solve(out, shape, tol=1e-12):
    class S:
        pass
    ps = S()
    ps.shape = shape
    ps.padded = lambda n: out[n]
    phi, r, nit, _ = oracle.cg(ps, max_iter=50000, tol=tol)
    assert r < tol
    return phi


def test_conservative_fluxes_balance_the_solved_system():
    """F = F* + a (pp_P - pp_nb) on every open liquid face: with pp the
    solution, the net outflow of every cell equals -(aP - sum a_nb) pp_P
    (compliance and acoustic terms), because aP pp_P - sum a_nb pp_nb = Su
    and Su is the net inflow of F*.  Checks the exports, the flux formula's
    sign and neighbour, and the periodic halo of the theta fluxes."""
    shape = (14, 24, 20)
    F, cfg = fields.make_fields(*shape, seed=8, two_phase=False, scrap=False)
    fluid = F["cell_type"] != 0
    F["alpha_l"] = np.asfortranarray(np.where(fluid, 1.0, 0.0))
    F["alpha_g"] = np.zeros_like(F["alpha_g"], order="F")
    F["outlet"] = np.zeros_like(F["outlet"])
    out, _, _, ex = oracle.pp_assemble(F, cfg, shape, exports=True)
    pp = _solve(out, shape)
    c = oracle.pp_correct(F, cfg, shape, 0.5, pp, ex)
    Fr, Ft, Fz = c["Fc_r"], c["Fc_th"], c["Fc_z"]
    net = (Fr[2:-2, 2:-2, 2:-2] - Fr[1:-3, 2:-2, 2:-2] + Ft[2:-2, 2:-2, 2:-2] - Ft[2:-2, 1:-3, 2:-2] +
           Fz[2:-2, 2:-2, 2:-2] - Fz[2:-2, 2:-2, 1:-3])
    a = {n: interior(out[n]) for n in NAMES}
    diag = a["aP"] - (a["aW"] + a["aE"] + a["aS"] + a["aN"] + a["aB"] + a["aT"])
    want = -diag * interior(pp)
    act = interior(fluid)
    scale = np.abs(Fz).max()
    assert np.abs(net[act] - want[act]).max() <= 1e-8 * scale
    assert np.abs(want[act]).max() > 1e-6 * scale          # the identity is not trivially 0 = 0


@pytest.mark.parametrize("axis", ["r", "theta", "z"])
def test_velocity_correction_of_a_linear_pp(axis):
    """pp linear along one axis (c*r, c*theta, c*z): the central difference
    is c (c/r in theta), so each active cell's velocity along that axis moves
    by -alpha*V/aP_u*c (c/r) (walls excepted in r and z) and the others do not
    move."""
    shape = (12, 18, 14)
    F, cfg = fields.make_fields(*shape, seed=9, masks=True)
    _, _, _, ex = oracle.pp_assemble(F, cfg, shape, exports=True)
    r, th, z = F["r"], F["theta"], F["z"]
    c = 123.0
    shp = F["p"].shape
    if axis == "r":
        pp = c * np.broadcast_to(r[:, None, None], shp)
    elif axis == "theta":
        pp = c * np.broadcast_to(th[None, :, None], shp)
    else:
        pp = c * np.broadcast_to(z[None, None, :], shp)
    pp = np.asfortranarray(pp)
    out = oracle.pp_correct(F, cfg, shape, 0.5, pp, ex)
    comp = {"r": 0, "theta": 1, "z": 2}[axis]
    ct = interior(F["cell_type"]) != 0
    act = {"l": (interior(F["liq_cont"]) != 0) & ct}
    ag = (interior(F["alpha_g"]) >= 1e-2) & (interior(F["pv_active"]) != 0)
    ag |= (interior(F["gas_reach"]) != 0) & (interior(F["alpha_g"]) < 1e-2) & ct
    act["g"] = ag & ct
    for ph in ("l", "g"):
        for q, name in enumerate(("ur", "uth", "uz")):
            du = interior(out[f"{name}_{ph}"]) - interior(F[f"{name}_{ph}"])
            if q != comp:
                assert np.abs(du).max() <= 1e-9 * c, (ph, name)
                continue
            apu = interior(F[{0: "aPr", 1: "aPth", 2: "aPz"}[q] + f"_{ph}"])
            want = -interior(F[f"alpha_{ph}"]) * interior(F["vol"]) / apu * c
            if q == 1:
                want = want / r[2:-2, None, None]
            sel = act[ph].copy()
            if q == 0:
                sel[[0, -1]] = False
            if q == 2:
                sel[:, :, [0, -1]] = False
            np.testing.assert_allclose(du[sel], want[sel], rtol=1e-9, atol=1e-12 * c)
            assert not du[~act[ph]].any()


def test_pressure_update_and_sealed_cells():
    shape = (12, 18, 14)
    F, cfg = fields.make_fields(*shape, seed=10)
    _, _, _, ex = oracle.pp_assemble(F, cfg, shape, exports=True)
    rng = np.random.default_rng(3)
    pp = np.asfortranarray(rng.normal(0, 5e5, F["p"].shape))
    out = oracle.pp_correct(F, cfg, shape, 0.7, pp, ex)
    upd = np.clip(F["p"] + 0.7 * pp, -2e6, 2e6)
    ct = F["cell_type"] != 0
    act = ((F["alpha_l"] >= 1e-2) | (F["alpha_g"] >= 1e-2)) & ct
    sealed = ct & ~act
    sealed[:2] = sealed[-2:] = False
    sealed[:, :2] = sealed[:, -2:] = False
    sealed[:, :, :2] = sealed[:, :, -2:] = False
    assert sealed.any()
    p = out["p"]
    inner = np.zeros_like(ct)
    inner[2:-2, 2:-2, 2:-2] = True
    keep = inner & ~sealed
    np.testing.assert_array_equal(p[keep], upd[keep])
    # every sealed cell sits within the range of its updated fluid neighbours, or is 0 without one
    for i, j, k in zip(*np.nonzero(sealed)):
        nb = [(i - 1, j, k), (i + 1, j, k), (i, j - 1, k), (i, j + 1, k), (i, j, k - 1), (i, j, k + 1)]
        vals = [upd[q] for q in nb if act[q]]
        if vals:
            assert min(vals) - 1e-9 <= p[i, j, k] <= max(vals) + 1e-9
        else:
            assert p[i, j, k] == 0.0
    nth = shape[1]
    np.testing.assert_array_equal(p[:, 1, :], p[:, nth + 1, :])
    np.testing.assert_array_equal(p[:, nth + 2, :], p[:, 2, :])
