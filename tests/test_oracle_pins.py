"""Pins the oracle functions that no property of tests/test_oracle.py or
tests/test_pp_oracle.py fixes by value: the true residual
(oracle_true_residual), the magnitude of the p' face coefficients, the
exact diagonal, and the gas compressibility source of the assembly
(oracle_pp_assemble_ex, src/mod_pressure_3d.f90:171-306,469-610).

Every expected value is built here from the inputs and the geometry, in
numpy, and every check is run twice: on the oracle as built, where it must
pass, and on the oracle rebuilt with a planted mistake (``oracle.planted``,
the ORACLE_PLANT hooks of the C sources), where it must fail.  The planted
mistakes are the plausible ones: a harmonic d_f, a dropped alpha_f average,
the wrong face area, a dropped compliance factor, the gas acoustic term
without 1/dt, c instead of c^2, no cap or no temperature floor on the
compressibility source, stale theta halos or dropped wall halos in the
residual.
"""
import math

import numpy as np
import pytest

import oracle
from eafp import fields, systems

NAMES = ("aW", "aE", "aS", "aN", "aB", "aT", "aP", "Su")
PP_COMPLIANCE = 1.0e-6      # src/mod_pressure_3d.f90:48
COMP_SRC_CAP = 0.2          # :54
P0_THERMO = 101325.0        # :63
C_SOUND_LIQ = 4000.0        # :67
T_MIN_GAS = 100.0           # src/mod_constants.f90:148


def interior(a):
    return a[2:-2, 2:-2, 2:-2]


class _Sys:
    """A system given by its padded arrays (the oracle's ``ps`` protocol)."""

    def __init__(self, shape, arrs):
        self.shape = shape
        self.nr, self.nth, self.nz = shape
        self._a = arrs

    def padded(self, n):
        return self._a[n]

    def interior(self, n):
        return interior(self._a[n])


# ---------------------------------------------------------------------------
# oracle_true_residual
# ---------------------------------------------------------------------------

def _residual_case(seed=0):
    """Production-like system with nonzero wall links (aW at i = 1, aE at
    i = nr, aB at k = 1, aT at k = nz), an x with nonzero r/z halos and
    garbage in its theta halo planes."""
    ps = systems.make_system(1)
    rng = np.random.default_rng(seed)
    arrs = {n: np.array(ps.padded(n), order="F") for n in NAMES}
    act = interior(arrs["aP"]) != 1.0
    for n, sl in (("aW", (2, slice(2, -2), slice(2, -2))), ("aE", (-3, slice(2, -2), slice(2, -2))),
                  ("aB", (slice(2, -2), slice(2, -2), 2)), ("aT", (slice(2, -2), slice(2, -2), -3))):
        w = rng.uniform(0.1, 1.0, arrs[n][sl].shape) * np.abs(arrs["aP"][sl]).max()
        arrs[n][sl] = np.where(arrs["aP"][sl] != 1.0, w, 0.0)
        arrs["aP"][sl] = arrs["aP"][sl] + arrs[n][sl]
    x = np.asfortranarray(rng.normal(size=arrs["aP"].shape))
    nth = ps.nth
    x[:, [0, 1, nth + 2, nth + 3], :] = rng.normal(1e3, 1.0, (x.shape[0], 4, x.shape[2]))   # stale theta halos
    assert act.any()
    return _Sys(ps.shape, arrs), x


def _residual_expected(S, x):
    """b - A x from scipy's CSR operator of the interior (theta periodic, wall
    links dropped), plus each wall link times x's r/z halo value."""
    A = oracle.csr(S)
    xi = interior(x)
    b = S.interior("Su")
    fold = np.zeros_like(xi)
    fold[0] += S.interior("aW")[0] * x[1, 2:-2, 2:-2]
    fold[-1] += S.interior("aE")[-1] * x[-2, 2:-2, 2:-2]
    fold[:, :, 0] += S.interior("aB")[:, :, 0] * x[2:-2, 2:-2, 1]
    fold[:, :, -1] += S.interior("aT")[:, :, -1] * x[2:-2, 2:-2, -2]
    r = b.ravel() - (A @ xi.ravel() - fold.ravel())
    return float(np.linalg.norm(r) / np.linalg.norm(b)), float(np.abs(r).sum() / np.abs(b).sum())


def _residual_ok(S, x):
    l2, l1 = oracle.true_residual(S, x)
    e2, e1 = _residual_expected(S, x)
    return math.isclose(l2, e2, rel_tol=1e-12) and math.isclose(l1, e1, rel_tol=1e-12)


def test_true_residual_matches_csr_with_wall_halos():
    S, x = _residual_case()
    assert _residual_ok(S, x)
    # the halo rule: theta halos are ignored (periodic source), r/z halos count
    x2 = np.array(x, order="F")
    oracle.theta_halo(S.shape, x2)
    assert oracle.true_residual(S, x2) == oracle.true_residual(S, x)
    x3 = np.array(x, order="F")
    x3[1] += 1.0
    assert oracle.true_residual(S, x3)[0] != oracle.true_residual(S, x)[0]


def test_true_residual_zero_at_the_solution():
    """x from scipy's direct solve of the gauge-fixed operator gives a
    residual at rounding level; a perturbed x does not."""
    import scipy.sparse.linalg as sla
    S, x = _residual_case(1)
    A = oracle.csr(S).tocsc()
    b = S.interior("Su").ravel()
    xs = sla.spsolve(A, b)
    xp = np.zeros_like(x)
    xp[2:-2, 2:-2, 2:-2] = xs.reshape(S.shape)
    l2, l1 = oracle.true_residual(S, np.asfortranarray(xp))
    assert l2 < 1e-9 and l1 < 1e-9
    xp[5, 5, 5] += 1e-3
    assert oracle.true_residual(S, np.asfortranarray(xp))[0] > 1e-6


@pytest.mark.parametrize("plant", [9, 10])
def test_true_residual_planted_mistakes_fail(plant):
    S, x = _residual_case()
    with oracle.planted(plant):
        assert not _residual_ok(S, x)


# ---------------------------------------------------------------------------
# Face coefficients of a single-phase mesh: closed form from the geometry
# ---------------------------------------------------------------------------

SHAPE = (12, 18, 14)


def _geometry(nr, nth, nz):
    """Cell geometry restated from the mesh definition (uniform faces from
    R_AXIS_MIN to R in r, 2 pi/nth in theta, 0..H in z; src/mod_mesh_3d.f90:
    70-156), not read from the field arrays."""
    rf = np.linspace(systems.R_AXIS_MIN, systems.R_SHELL, nr + 1)
    rc = 0.5 * (rf[:-1] + rf[1:])
    dth = 2.0 * math.pi / nth
    dz = systems.H_TOTAL / nz
    ring = 0.5 * (rf[1:] ** 2 - rf[:-1] ** 2)
    V = np.broadcast_to((ring * dth * dz)[:, None, None], (nr, nth, nz))
    return rf, rc, dth, dz, ring, V


def _single_phase_case(seed=11):
    """Liquid only (the gas is not in the Poisson system): alpha_l random in
    [0.3, 0.9] in every fluid cell, the three aP_u of each cell alternating
    between 3 and 50 in a checkerboard, no outlets."""
    nr, nth, nz = SHAPE
    F, cfg = fields.make_fields(nr, nth, nz, seed=seed, two_phase=False, scrap=False)
    rng = np.random.default_rng(seed)
    fluid = F["cell_type"] != 0
    F["alpha_l"] = np.where(fluid, rng.uniform(0.3, 0.9, fluid.shape), 0.0)
    ii, jj, kk = np.indices(fluid.shape)
    cb = np.where((ii + jj + kk) % 2 == 0, 3.0, 50.0)
    for n in ("aPr_l", "aPth_l", "aPz_l"):
        F[n] = cb.copy()
    for n in ("alpha_l", "aPr_l", "aPth_l", "aPz_l"):
        F[n] = fields.theta_halo(np.asfortranarray(F[n]))
    F["outlet"] = np.zeros_like(F["outlet"])
    return F, cfg


def _face_expected(F):
    """Closed form of each face coefficient: a_f = alpha_f * d_f * A_f / delta_f
    with alpha_f the arithmetic face mean of alpha and d_f the arithmetic face
    mean of alpha V / aP_u (src/mod_pressure_3d.f90:496-505 and the five other
    faces), areas and distances from the cell geometry; 0 where either side
    is not fluid."""
    nr, nth, nz = SHAPE
    rf, rc, dth, dz, ring, V = _geometry(nr, nth, nz)
    al = interior(F["alpha_l"])
    fl = interior(F["cell_type"]) != 0
    out = {}
    q = al * V / interior(F["aPr_l"])
    aW = np.zeros(SHAPE)
    aW[1:] = 0.5 * (al[1:] + al[:-1]) * 0.5 * (q[1:] + q[:-1]) * (rf[1:-1] * dth * dz)[:, None, None] / \
        (rc[1:] - rc[:-1])[:, None, None]
    aW[1:] *= fl[1:] & fl[:-1]
    out["aW"] = aW
    aE = np.zeros(SHAPE)
    aE[:-1] = aW[1:]
    out["aE"] = aE
    q = al * V / interior(F["aPth_l"])
    qs, als, fls = np.roll(q, 1, axis=1), np.roll(al, 1, axis=1), np.roll(fl, 1, axis=1)
    dr = (rf[1:] - rf[:-1])[:, None, None]
    out["aS"] = 0.5 * (al + als) * 0.5 * (q + qs) * dr * dz / (rc[:, None, None] * dth) * (fl & fls)
    out["aN"] = np.roll(out["aS"], -1, axis=1)
    q = al * V / interior(F["aPz_l"])
    aB = np.zeros(SHAPE)
    aB[:, :, 1:] = 0.5 * (al[:, :, 1:] + al[:, :, :-1]) * 0.5 * (q[:, :, 1:] + q[:, :, :-1]) * \
        (ring * dth)[:, None, None] / dz
    aB[:, :, 1:] *= fl[:, :, 1:] & fl[:, :, :-1]
    out["aB"] = aB
    aT = np.zeros(SHAPE)
    aT[:, :, :-1] = aB[:, :, 1:]
    out["aT"] = aT
    return out


def _faces_ok(F, cfg):
    out, _, _ = oracle.pp_assemble(F, cfg, SHAPE)
    want = _face_expected(F)
    ok = True
    for n in ("aW", "aE", "aS", "aN", "aB", "aT"):
        got = interior(out[n])
        ok &= bool(np.allclose(got, want[n], rtol=1e-12, atol=0.0))
    return ok


def test_face_coefficients_closed_form():
    F, cfg = _single_phase_case()
    out, _, _ = oracle.pp_assemble(F, cfg, SHAPE)
    want = _face_expected(F)
    for n in ("aW", "aE", "aS", "aN", "aB", "aT"):
        assert np.count_nonzero(want[n]) > 0.3 * want[n].size, n
        np.testing.assert_allclose(interior(out[n]), want[n], rtol=1e-12, atol=0.0, err_msg=n)


@pytest.mark.parametrize("plant", [1, 2, 3])
def test_face_coefficients_planted_mistakes_fail(plant):
    F, cfg = _single_phase_case()
    with oracle.planted(plant):
        assert not _faces_ok(F, cfg)


# ---------------------------------------------------------------------------
# The exact diagonal (src/mod_pressure_3d.f90:261-279)
# ---------------------------------------------------------------------------

def _two_phase_case(seed=12):
    nr, nth, nz = SHAPE
    F, cfg = fields.make_fields(nr, nth, nz, seed=seed, two_phase=True, scrap=True)
    F["outlet"] = np.zeros_like(F["outlet"])
    return F, cfg


def _diag_expected(F, cfg, out):
    """aP = (1 + PP_COMPLIANCE) sum(a_nb) + [gas active] alpha_g V/(P0 dt)
    + [liquid active] alpha_l V/(rho_l c^2 dt) on rows with an active phase,
    1 on the others.  No link of these cases is removed by the boundary
    conditions (walls and masked cells are never linked, no outlets), so the
    returned links are the ones the diagonal summed."""
    a = {n: interior(out[n]) for n in NAMES}
    nb = a["aW"] + a["aE"] + a["aS"] + a["aN"] + a["aB"] + a["aT"]
    fl = interior(F["cell_type"]) != 0
    al, ag = interior(F["alpha_l"]), interior(F["alpha_g"])
    act_l = al >= 1e-2
    act_g = ag >= 1e-2
    V = interior(F["vol"])
    dt = cfg["dt"]
    two = bool(cfg["gas_compressibility"] and cfg["solve_multiphase"])
    gas = np.where(act_g & two, ag * V / (P0_THERMO * dt), 0.0)
    liq = np.where(act_l, al * V / (interior(F["rho_l"]) * C_SOUND_LIQ ** 2 * dt), 0.0)
    want = (1.0 + PP_COMPLIANCE) * nb + gas + liq
    rows = fl & (act_l | (act_g & bool(cfg["gas_in_poisson"] and cfg["solve_multiphase"])) | act_g)
    want = np.where(rows, want, 1.0)
    return want, rows, gas, liq


def _diag_ok(F, cfg):
    out, _, _ = oracle.pp_assemble(F, cfg, SHAPE)
    want, _, _, _ = _diag_expected(F, cfg, out)
    return bool(np.allclose(interior(out["aP"]), want, rtol=1e-13, atol=0.0))


@pytest.mark.parametrize("case", ["single_phase", "two_phase"])
def test_diagonal_exact(case):
    F, cfg = _single_phase_case() if case == "single_phase" else _two_phase_case()
    out, _, _ = oracle.pp_assemble(F, cfg, SHAPE)
    want, rows, gas, liq = _diag_expected(F, cfg, out)
    assert rows.sum() > 0.3 * rows.size
    assert np.count_nonzero(liq) > 0
    if case == "two_phase":
        assert np.count_nonzero(gas) > 0.2 * rows.sum()
    np.testing.assert_allclose(interior(out["aP"]), want, rtol=1e-13, atol=0.0)


@pytest.mark.parametrize("plant,case", [(4, "single_phase"), (5, "two_phase"), (6, "single_phase"),
                                        (6, "two_phase")])
def test_diagonal_planted_mistakes_fail(plant, case):
    F, cfg = _single_phase_case() if case == "single_phase" else _two_phase_case()
    with oracle.planted(plant):
        assert not _diag_ok(F, cfg)


# ---------------------------------------------------------------------------
# Gas compressibility source (src/mod_pressure_3d.f90:226-246)
# ---------------------------------------------------------------------------

RATIOS = np.array([0.5, 0.85, 0.95, 1.0, 1.05, 1.15, 1.5, 3.0])


def _comp_case(seed=13):
    """The rest state (no velocity, uniform phases, pressure linear in z: every
    face flux vanishes) with the gas heated or cooled at constant pressure by
    a per-cell ratio q = T/T_old: rho_g = rho_gas T_ambient / T with
    T = q max(T_old, T_MIN_GAS); a quarter of the cells have T_old below
    T_MIN_GAS."""
    shape = SHAPE
    F, cfg = fields.make_fields(*shape, seed=seed, two_phase=True, scrap=False)
    rng = np.random.default_rng(seed)
    fluid = F["cell_type"] != 0
    for n in ("ur_l", "uth_l", "uz_l", "ur_g", "uth_g", "uz_g"):
        F[n] = np.zeros_like(F[n], order="F")
    F["alpha_l"] = np.asfortranarray(np.where(fluid, 0.6, 0.0))
    F["alpha_g"] = np.asfortranarray(np.where(fluid, 0.4, 0.0))
    for n, v in (("rho_l", 7000.0), ("T_l", 1850.0)):
        F[n] = np.full_like(F[n], v, order="F")
    F["p"] = np.asfortranarray(np.broadcast_to(-3000.0 * F["z"][None, None, :], F["p"].shape).copy())
    F["outlet"] = np.zeros_like(F["outlet"])
    q = RATIOS[rng.integers(0, RATIOS.size, fluid.shape)]
    T_old = np.where(rng.random(fluid.shape) < 0.25, rng.uniform(20.0, 90.0, fluid.shape),
                     rng.uniform(300.0, 3000.0, fluid.shape))
    T = q * np.maximum(T_old, T_MIN_GAS)
    F["T_old_g"] = fields.theta_halo(np.asfortranarray(T_old))
    F["rho_g"] = fields.theta_halo(np.asfortranarray(cfg["rho_gas"] * cfg["T_ambient"] / T))
    F["_q"] = q
    return F, cfg


def _comp_expected(F, cfg):
    """Su = alpha_g V / dt * clamp(q - 1, -COMP_SRC_CAP, COMP_SRC_CAP): the
    volume form (alpha_g/rho_g)(rho(T) - rho(T_old)) V/dt of an ideal gas at
    constant pressure is alpha_g V/dt (1 - T/T_old), subtracted from Su."""
    fl = interior(F["cell_type"]) != 0
    m = interior(F["alpha_g"]) * interior(F["vol"]) / cfg["dt"]
    want = m * np.clip(interior(F["_q"]) - 1.0, -COMP_SRC_CAP, COMP_SRC_CAP)
    return np.where(fl, want, 0.0), m


def _comp_ok(F, cfg):
    G = {k: v for k, v in F.items() if not k.startswith("_")}
    out, _, _ = oracle.pp_assemble(G, cfg, SHAPE)
    want, m = _comp_expected(F, cfg)
    return bool(np.abs(interior(out["Su"]) - want).max() <= 1e-9 * m.max())


def test_compressibility_source_closed_form():
    F, cfg = _comp_case()
    G = {k: v for k, v in F.items() if not k.startswith("_")}
    out, _, _ = oracle.pp_assemble(G, cfg, SHAPE)
    want, m = _comp_expected(F, cfg)
    q = interior(F["_q"])
    fl = interior(F["cell_type"]) != 0
    assert (fl & (np.abs(q - 1.0) > COMP_SRC_CAP)).any()          # capped cells present
    assert (fl & (np.abs(q - 1.0) < COMP_SRC_CAP) & (q != 1.0)).any()
    assert (fl & (interior(F["T_old_g"]) < T_MIN_GAS)).any()
    np.testing.assert_allclose(interior(out["Su"]), want, rtol=0, atol=1e-9 * m.max())


@pytest.mark.parametrize("plant", [7, 8])
def test_compressibility_planted_mistakes_fail(plant):
    F, cfg = _comp_case()
    with oracle.planted(plant):
        assert not _comp_ok(F, cfg)


# ---------------------------------------------------------------------------
# The Rhie-Chow pressure term of u_f (src/mod_pressure_3d.f90:496-530)
# ---------------------------------------------------------------------------

def _alternating_case():
    """The single-phase case at rest with p = 1e5 + 50 (-1)^i: the central
    cell gradients of an r-alternating pressure vanish, so on the faces of a
    cell whose two r-neighbours on each side are fluid, u_f reduces to
    -d_f (p_N - p_P)/delta and each r face adds -a_f (p_P - p_N) to Su;
    theta and z faces see no pressure difference."""
    F, cfg = _single_phase_case()
    for n in ("ur_l", "uth_l", "uz_l", "ur_g", "uth_g", "uz_g"):
        F[n] = np.zeros_like(F[n], order="F")
    ii = np.indices(F["p"].shape)[0] - 1
    F["p"] = np.asfortranarray(1e5 + 50.0 * np.where(ii % 2 == 0, 1.0, -1.0))
    return F, cfg


def _rc_check(F, cfg):
    out, _, _ = oracle.pp_assemble(F, cfg, SHAPE)
    nr = SHAPE[0]
    fl = (F["cell_type"] != 0) & (F["alpha_l"] >= 1e-2)
    sel = np.zeros(SHAPE, dtype=bool)
    for i in range(3, nr - 1):          # Fortran i; numpy index i + 1 of the padded arrays
        sel[i - 1] = fl[i - 1:i + 4].all(axis=0)[2:-2, 2:-2]
    p = interior(F["p"])
    dp = p - np.roll(p, 1, axis=0)      # p_P - p_W (p_P - p_E is the same for an alternating p)
    want = -(interior(out["aW"]) + interior(out["aE"])) * dp
    got = interior(out["Su"])
    return sel, got, want


def test_rhie_chow_pressure_term():
    F, cfg = _alternating_case()
    sel, got, want = _rc_check(F, cfg)
    assert sel.sum() > 100
    scale = np.abs(want[sel]).max()
    assert scale > 0
    np.testing.assert_allclose(got[sel], want[sel], rtol=0, atol=1e-10 * scale)


def test_rhie_chow_planted_mistake_fails():
    F, cfg = _alternating_case()
    with oracle.planted(11):
        sel, got, want = _rc_check(F, cfg)
    assert np.abs(got[sel] - want[sel]).max() > 1e-3 * np.abs(want[sel]).max()


def test_unplanted_rebuild_matches_the_oracle():
    """ORACLE_PLANT=0 is the oracle as built: the same bits."""
    F, cfg = _two_phase_case()
    out, res, flux = oracle.pp_assemble(F, cfg, SHAPE)
    with oracle.planted(0):
        out0, res0, flux0 = oracle.pp_assemble(F, cfg, SHAPE)
    for n in NAMES:
        np.testing.assert_array_equal(out0[n], out[n])
    assert (res0, flux0) == (res, flux)
