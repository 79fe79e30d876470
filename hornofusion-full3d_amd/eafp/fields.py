"""Seeded synthetic inputs of the p' assembly (the fields that
``solve_pressure_correction``, src/mod_pressure_3d.f90:88-318, reads).

The mesh follows src/mod_mesh_3d.f90 on one serial rank:
- r and z faces are uniform and clamped at the walls in the halos (:79-103):
  the halo cell centres sit on the wall and the halo widths are 0;
- theta faces are unrolled, 2*pi*j/ntheta for every j (:87-98);
- r, theta, z, vol, Ar, Ath and Az follow :105-156;
- cell_type is 0 below the parabolic bowl and in every r/z halo (:159-178, :226-243);
- the outlet mask is the top layer of the electrode columns (src/mod_mesh_3d.f90:279-307,
  src/mod_boundary_3d.f90:203-214).

The phase fields describe a liquid-steel bath (rho ~ 7000, alpha_l = 1 below a jittered surface, one transition
layer), gas above it, and an optional scrap band in which alpha_l + alpha_g < 1. They carry:
- small random velocities;
- log-uniform momentum diagonals aP_u;
- a hydrostatic pressure with noise;
- gas temperatures for the compressibility term.

The optional workspace masks (ws_liq_cont, ws_pv_active, ws_gas_reach, ws_liq_room) are drawn at random when asked
for.

Every 3-D array is float64 (int32 for masks) in the reference's padded Fortran layout: shape (nr+4, nth+4, nz+4),
order 'F', Fortran index i at numpy index i+1. The theta halos hold the periodic copies. The 1-D coordinates have
bounds (-1:n+2).
"""
from __future__ import annotations

import math

import numpy as np

from .systems import H_BOWL, H_TOTAL, N_ELECTRODES, R_AXIS_MIN, R_BOWL, R_ELEC, R_PCD, R_SHELL, SEED_BASE, TWO_PI

FIELD3 = ("vol", "Ar", "Ath", "Az",
          "alpha_l", "ur_l", "uth_l", "uz_l", "aPr_l", "aPth_l", "aPz_l", "rho_l", "T_l",
          "alpha_g", "ur_g", "uth_g", "uz_g", "aPr_g", "aPth_g", "aPz_g", "rho_g", "T_old_g", "p")
MASKS = ("liq_cont", "pv_active", "gas_reach")


def _faces(n, lo, hi):
    """Faces at Fortran indices -2 .. n+2, clamped to [lo, hi] beyond the walls."""
    f = lo + (hi - lo) * np.arange(n + 1, dtype=np.float64) / n
    idx = np.clip(np.arange(-2, n + 3), 0, n)
    return f[idx]


def theta_halo(a):
    """The serial periodic theta copy (src/mod_mpi_topology.f90:303-310) on a padded array."""
    nth = a.shape[1] - 4
    a[:, 1, :] = a[:, nth + 1, :]      # j = 0      <- j = nth
    a[:, 0, :] = a[:, nth, :]          # j = -1     <- j = nth-1
    a[:, nth + 2, :] = a[:, 2, :]      # j = nth+1  <- j = 1
    a[:, nth + 3, :] = a[:, 3, :]      # j = nth+2  <- j = 2
    return a


def make_fields(nr, nth, nz, seed=0, two_phase=True, masks=False, scrap=True, H=H_TOTAL, R=R_SHELL):
    """Returns (fields, config): fields maps every input name to its array."""
    rng = np.random.default_rng(SEED_BASE + 7919 * seed + 13 * nr + nz)
    rf = _faces(nr, R_AXIS_MIN, R)                           # index i+2
    zf = _faces(nz, 0.0, H)
    thf = TWO_PI * np.arange(-2, nth + 3, dtype=np.float64) / nth
    r = 0.5 * (rf[:-1] + rf[1:])                             # index i+1, i = -1 .. nr+2
    dr = rf[1:] - rf[:-1]
    th = 0.5 * (thf[:-1] + thf[1:])
    dth = thf[1:] - thf[:-1]
    z = 0.5 * (zf[:-1] + zf[1:])
    dz = zf[1:] - zf[:-1]
    shp = (nr + 4, nth + 4, nz + 4)
    R3 = np.broadcast_to(r[:, None, None], shp)
    F = {}
    F["r"], F["theta"], F["z"] = r.copy(), th.copy(), z.copy()
    ring = 0.5 * (rf[1:] ** 2 - rf[:-1] ** 2)
    F["vol"] = np.asfortranarray(ring[:, None, None] * dth[None, :, None] * dz[None, None, :])
    F["Ar"] = np.asfortranarray(rf[1:, None, None] * dth[None, :, None] * dz[None, None, :])
    F["Ath"] = np.asfortranarray(np.broadcast_to(dr[:, None, None] * dz[None, None, :], shp))
    F["Az"] = np.asfortranarray(np.broadcast_to(ring[:, None, None] * dth[None, :, None], shp))
    zb = np.where(r <= R_BOWL, H_BOWL * (r / R_BOWL) ** 2, H_BOWL)
    ct = np.ones(shp, dtype=np.int32)
    ct[np.broadcast_to((z[None, None, :] < zb[:, None, None]) | (R3 > R), shp)] = 0
    ct[:2, :, :] = 0
    ct[-2:, :, :] = 0
    ct[:, :, :2] = 0
    ct[:, :, -2:] = 0
    F["cell_type"] = np.asfortranarray(ct)
    x = r[:, None] * np.cos(th[None, :])
    y = r[:, None] * np.sin(th[None, :])
    out = np.zeros((nr + 4, nth + 4), dtype=np.int32)
    for e in range(N_ELECTRODES):
        the = TWO_PI * e / N_ELECTRODES
        out |= (np.sqrt((x - R_PCD * math.cos(the)) ** 2 + (y - R_PCD * math.sin(the)) ** 2) <= R_ELEC).astype(np.int32)
    F["outlet"] = np.asfortranarray(out)

    # phases
    fluid = ct != 0
    z3 = np.broadcast_to(z[None, None, :], shp)
    zs = 0.9 + rng.uniform(-0.05, 0.05) + 0.02 * np.sin(th)[None, :, None] * np.ones(shp)
    al = np.clip((zs - (z3 - 0.5 * dz[None, None, :])) / np.maximum(dz[None, None, :], 1e-30), 0.0, 1.0)
    solid = np.zeros(shp)
    if scrap:
        band = (z3 > 1.2) & (z3 < 2.0) & (R3 > 0.8) & (R3 < 2.8)
        solid = np.where(band, rng.uniform(0.3, 0.999, shp), 0.0)
    al = al * (1.0 - solid)
    ag = np.clip(1.0 - al - solid, 0.0, 1.0)
    # a sprinkle of sub-cutoff liquid and gas fractions
    trace = rng.random(shp) < 0.03
    al = np.where(trace & (al < 0.5), rng.uniform(0, 0.02, shp), al)
    ag = np.where(trace & (ag < 0.5), rng.uniform(0, 0.02, shp), ag)
    al[~fluid] = 0.0
    ag[~fluid] = 0.0
    F["alpha_l"], F["alpha_g"] = al, ag
    for ph, su in (("l", 0.05), ("g", 0.5)):
        for c in ("ur", "uth", "uz"):
            F[f"{c}_{ph}"] = rng.normal(0.0, su, shp)
        for c in ("aPr", "aPth", "aPz"):
            F[f"{c}_{ph}"] = np.exp(rng.uniform(math.log(1.0), math.log(1e4), shp))
    F["rho_l"] = 7000.0 * (1.0 + 0.01 * rng.normal(size=shp))
    F["T_l"] = 1850.0 + 20.0 * rng.normal(size=shp)
    F["rho_g"] = 0.3 * (1.0 + 0.2 * rng.random(shp))
    F["T_old_g"] = 1500.0 + 1500.0 * rng.random(shp)
    rho_mix = np.where(z3 < zs, 7000.0, 0.3)
    F["p"] = rho_mix * 9.81 * (zs - z3) + 10.0 * rng.normal(size=shp)
    if masks:
        F["liq_cont"] = ((al >= 1e-2) & (rng.random(shp) > 0.05)).astype(np.int32)
        F["pv_active"] = (rng.random(shp) > 0.02).astype(np.int32)
        F["gas_reach"] = (rng.random(shp) > 0.5).astype(np.int32)
        room = rng.random(shp)
        F["liq_room"] = np.where(rng.random(shp) < 0.1, rng.uniform(0, 1e-4, shp), room)
    for k, v in list(F.items()):
        if isinstance(v, np.ndarray) and v.ndim == 3:
            F[k] = theta_halo(np.asfortranarray(v))
    cfg = {"dt": 1e-3, "rho_gas": 1.2, "T_ambient": 300.0, "beta_expansion": 1e-4,
           "gas_in_poisson": int(two_phase), "solve_multiphase": int(two_phase), "gas_compressibility": int(two_phase)}
    return F, cfg
