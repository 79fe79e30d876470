"""Seeded synthetic p' (pressure-correction) systems on the EAF cylinder.

Each system is the 7-point cylindrical (r, theta, z) Poisson operator that
EAF3D assembles before its conjugate-gradient call, built on the reference's
geometry conventions:

* mesh: r faces uniform from R_AXIS_MIN = 0.02 m to R (the axis is a
  zero-flux inner wall, ``src/mod_constants.f90:151``,
  ``src/mod_mesh_3d.f90:70-71``), theta faces 2*pi*j/ntheta unrolled across
  the seam (``src/mod_mesh_3d.f90:95-98``), uniform z faces; ``vol``, ``Ar``,
  ``Ath``, ``Az`` as in ``src/mod_mesh_3d.f90:125-156``;
* crucible mask ``z_c < z_bowl(r_c)`` with a parabolic bowl
  (``src/mod_mesh_3d.f90:159-178``) and three electrode columns on the pitch
  circle (``src/mod_mesh_3d.f90:279-307``), masked over a z range;
* masked rows are identity rows (aP = 1, a_nb = 0, Su = 0) and links into
  them are zero (``src/mod_boundary_3d.f90:219-225,256-275``); walls are
  Neumann (``src/mod_boundary_3d.f90:228-230``);
* face coefficients ``a_f = rho_f * A_f * 0.5*(V_P/aPu_P + V_N/aPu_N) / delta_f``,
  the form of ``add_phase_contribution`` (``src/mod_pressure_3d.f90:496-505``)
  with a density weight; the same value is written from both sides, so the
  matrix is exactly symmetric;
* ``aP = sum(a_nb)`` (pure Neumann), so each connected active component is
  singular with the constant null vector; the RHS is N(0, 1e-3) projected to
  zero sum per component (FP64).

Arrays are produced in the reference's storage: Fortran order, bounds
``(-1:n+2)`` in each axis (``src/mod_workspace.f90:124-125``), i.e. numpy
arrays of shape ``(nr+4, nth+4, nz+4)`` with ``order='F'`` where Fortran index
``i`` sits at numpy index ``i+1``.  Halo entries are zero.

The five named workloads follow the p' benchmark definitions:
  1  12x24x16 gas only (rho = 1.2)
  2  60x120x84, liquid (7000) below 0.9 m, gas 0.3 above
  3  ensemble of uneven meshes 18..35 x 36..66 (step 6) x 24..50
  4  120x240x168, liquid pool + scrap band (Ergun factor 1e-2*(1-alpha_s))
  5  ensemble of 60x120x84 campaign variants (immersion, slag foam, jitter)
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

R_AXIS_MIN = 0.02          # src/mod_constants.f90:151
TWO_PI = 2.0 * math.pi
R_SHELL = 3.8
H_TOTAL = 4.5
R_BOWL = 2.5
H_BOWL = 0.6
R_PCD = 0.85
R_ELEC = 0.30
N_ELECTRODES = 3
SEED_BASE = 0xEAF3D

COEF_NAMES = ("aW", "aE", "aS", "aN", "aB", "aT", "aP")


@dataclasses.dataclass
class Geometry:
    nr: int
    nth: int
    nz: int
    rf: np.ndarray      # (nr+1,) radial faces
    r: np.ndarray       # (nr,) cell centres
    dr: np.ndarray
    dth: float
    theta: np.ndarray   # (nth,)
    dz: float
    z: np.ndarray       # (nz,)


@dataclasses.dataclass
class PSystem:
    """One assembled p' system in the reference's padded Fortran layout."""
    nr: int
    nth: int
    nz: int
    coef: dict          # name -> (nr+4, nth+4, nz+4) float64, order='F'
    Su: np.ndarray      # (nr+4, nth+4, nz+4) float64, order='F'
    active: np.ndarray  # (nr, nth, nz) bool, z fastest
    labels: np.ndarray  # (nr, nth, nz) int32 component id, -1 on masked cells
    ncomp: int
    meta: dict

    @property
    def shape(self):
        return (self.nr, self.nth, self.nz)

    @property
    def ncells(self):
        return self.nr * self.nth * self.nz

    def interior(self, name: str) -> np.ndarray:
        """Interior block as a C-contiguous (nr, nth, nz) array (z fastest)."""
        a = self.Su if name == "Su" else self.coef[name]
        return np.ascontiguousarray(a[2:-2, 2:-2, 2:-2])

    def padded(self, name: str) -> np.ndarray:
        return self.Su if name == "Su" else self.coef[name]


def make_geometry(nr: int, nth: int, nz: int, R: float = R_SHELL, H: float = H_TOTAL) -> Geometry:
    rf = R_AXIS_MIN + (R - R_AXIS_MIN) * np.arange(nr + 1, dtype=np.float64) / nr
    r = 0.5 * (rf[:-1] + rf[1:])
    dr = rf[1:] - rf[:-1]
    dth = TWO_PI / nth
    theta = (np.arange(nth, dtype=np.float64) + 0.5) * dth
    dz = H / nz
    z = (np.arange(nz, dtype=np.float64) + 0.5) * dz
    return Geometry(nr, nth, nz, rf, r, dr, dth, theta, dz, z)


def crucible_mask(g: Geometry) -> np.ndarray:
    """True where the cell lies below the refractory bowl (inactive)."""
    zb = np.where(g.r <= R_BOWL, H_BOWL * (g.r / R_BOWL) ** 2, H_BOWL)
    return np.broadcast_to(g.z[None, None, :] < zb[:, None, None], (g.nr, g.nth, g.nz))


def electrode_columns(g: Geometry) -> np.ndarray:
    """(nr, nth) True where the cell centre lies inside an electrode circle."""
    x = g.r[:, None] * np.cos(g.theta[None, :])
    y = g.r[:, None] * np.sin(g.theta[None, :])
    col = np.zeros((g.nr, g.nth), dtype=bool)
    for e in range(N_ELECTRODES):
        th = TWO_PI * e / N_ELECTRODES
        d = np.sqrt((x - R_PCD * math.cos(th)) ** 2 + (y - R_PCD * math.sin(th)) ** 2)
        col |= d <= R_ELEC
    return col


def _face_coefficients(g: Geometry, active, rho, apu, factor=None):
    """Face coefficients (a_r, a_th, a_z) on the +faces of every cell.

    a_r[i,j,k] couples (i,j,k)-(i+1,j,k) (zero for i = nr-1, the shell wall),
    a_th[i,j,k] couples (i,j,k)-(i,(j+1)%nth,k) (periodic),
    a_z[i,j,k] couples (i,j,k)-(i,j,k+1) (zero for k = nz-1, the roof).
    """
    nr, nth, nz = g.nr, g.nth, g.nz
    vol = (0.5 * (g.rf[1:] ** 2 - g.rf[:-1] ** 2))[:, None, None] * g.dth * g.dz
    vol = np.broadcast_to(vol, (nr, nth, nz))
    apr, apt, apz = apu
    fac = factor if factor is not None else None

    def face(vP, vN, aP_, aN_, rP, rN, A, delta, fP=None, fN=None):
        d = 0.5 * (vP / aP_ + vN / aN_)
        rho_f = 0.5 * (rP + rN)
        a = rho_f * A * d / delta
        if fP is not None:
            a = a * (0.5 * (fP + fN))
        return a

    # r faces
    a_r = np.zeros((nr, nth, nz))
    A = (g.rf[1:-1] * g.dth * g.dz)[:, None, None]
    delta = (g.r[1:] - g.r[:-1])[:, None, None]
    a_r[:-1] = face(vol[:-1], vol[1:], apr[:-1], apr[1:], rho[:-1], rho[1:], A, delta,
                    None if fac is None else fac[:-1], None if fac is None else fac[1:])
    link = active[:-1] & active[1:]
    a_r[:-1] *= link
    # theta faces (periodic)
    a_t = np.zeros((nr, nth, nz))
    A = (g.dr * g.dz)[:, None, None]
    delta = (g.r * g.dth)[:, None, None]
    rolls = lambda x: np.roll(x, -1, axis=1)
    a_t[:] = face(vol, rolls(vol), apt, rolls(apt), rho, rolls(rho), A, delta,
                  None if fac is None else fac, None if fac is None else rolls(fac))
    a_t *= active & rolls(active)
    # z faces
    a_z = np.zeros((nr, nth, nz))
    A = (0.5 * (g.rf[1:] ** 2 - g.rf[:-1] ** 2) * g.dth)[:, None, None]
    a_z[:, :, :-1] = face(vol[:, :, :-1], vol[:, :, 1:], apz[:, :, :-1], apz[:, :, 1:],
                          rho[:, :, :-1], rho[:, :, 1:], A, g.dz,
                          None if fac is None else fac[:, :, :-1], None if fac is None else fac[:, :, 1:])
    a_z[:, :, :-1] *= active[:, :, :-1] & active[:, :, 1:]
    return a_r, a_t, a_z


def _components(active, a_r, a_t, a_z):
    """Connected components of the active-cell link graph (scipy csgraph)."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components
    nr, nth, nz = active.shape
    idx = np.arange(nr * nth * nz).reshape(nr, nth, nz)
    rows, cols = [], []
    m = a_r > 0
    rows.append(idx[m]); cols.append(np.roll(idx, -1, axis=0)[m])
    m = a_t > 0
    rows.append(idx[m]); cols.append(np.roll(idx, -1, axis=1)[m])
    m = a_z > 0
    rows.append(idx[m]); cols.append(np.roll(idx, -1, axis=2)[m])
    rows = np.concatenate(rows); cols = np.concatenate(cols)
    n = nr * nth * nz
    G = coo_matrix((np.ones(rows.size, dtype=np.int8), (rows, cols)), shape=(n, n)).tocsr()
    ncomp, lab = connected_components(G, directed=False)
    lab = lab.reshape(nr, nth, nz).astype(np.int32)
    # relabel: masked cells -> -1, active components 0..c-1
    lab = np.where(active, lab, -1)
    uniq = np.unique(lab[active])
    remap = -np.ones(ncomp, dtype=np.int32)
    remap[uniq] = np.arange(uniq.size, dtype=np.int32)
    lab = np.where(active, remap[np.maximum(lab, 0)], -1).astype(np.int32)
    return int(uniq.size), lab


def project_zero_mean(b: np.ndarray, labels: np.ndarray, ncomp: int) -> np.ndarray:
    """Subtract the per-component mean of ``b`` (FP64); masked cells -> 0."""
    out = np.zeros_like(b, dtype=np.float64)
    act = labels >= 0
    lab = labels[act]
    vals = b[act].astype(np.float64)
    sums = np.bincount(lab, weights=vals, minlength=ncomp)
    cnt = np.bincount(lab, minlength=ncomp)
    out[act] = vals - (sums / np.maximum(cnt, 1))[lab]
    return out


def assemble(g: Geometry, active, rho, apu, rng, factor=None, meta=None) -> PSystem:
    """Assemble a pure-Neumann p' system from per-cell fields."""
    nr, nth, nz = g.nr, g.nth, g.nz
    active = np.array(active, dtype=bool)
    a_r, a_t, a_z = _face_coefficients(g, active, rho, apu, factor)
    aE = a_r
    aW = np.zeros_like(a_r); aW[1:] = a_r[:-1]
    aN = a_t
    aS = np.roll(a_t, 1, axis=1)
    aT = a_z
    aB = np.zeros_like(a_z); aB[:, :, 1:] = a_z[:, :, :-1]
    aP = aW + aE + aS + aN + aB + aT
    # active cells with no link at all would give an empty row: treat them as
    # masked identity rows, like cells without any active phase
    isolated = active & (aP == 0.0)
    active = active & ~isolated
    aP = np.where(active, aP, 1.0)
    ncomp, labels = _components(active, a_r, a_t, a_z)
    b = rng.normal(0.0, 1e-3, size=(nr, nth, nz))
    b = project_zero_mean(b, labels, ncomp)

    def pad(x):
        p = np.zeros((nr + 4, nth + 4, nz + 4), order="F")
        p[2:-2, 2:-2, 2:-2] = x
        return p

    coef = {"aW": pad(aW), "aE": pad(aE), "aS": pad(aS), "aN": pad(aN),
            "aB": pad(aB), "aT": pad(aT), "aP": pad(aP)}
    return PSystem(nr, nth, nz, coef, pad(b), active, labels, ncomp, meta or {})


def _log_uniform(rng, lo, hi, shape):
    return 10.0 ** rng.uniform(math.log10(lo), math.log10(hi), size=shape)


def _apu(rng, shape, lo=1e-3, hi=1e2):
    """Momentum-diagonal fields a_P^u in r, theta, z (log-uniform)."""
    return tuple(_log_uniform(rng, lo, hi, shape) for _ in range(3))


def system_seed(config: int, index: int = 0) -> int:
    return SEED_BASE + 1000 * config + index


def make_system(config: int, index: int = 0, shape=None) -> PSystem:
    """Build system ``index`` of benchmark workload ``config`` (1..5).

    ``shape`` overrides the mesh (used by tests for scaled-down variants of a
    workload; the fields and masks scale with the geometry).
    """
    rng = np.random.default_rng(system_seed(config, index))
    meta = {"config": config, "index": index}
    if config == 1:
        nr, nth, nz = shape or (12, 24, 16)
        g = make_geometry(nr, nth, nz)
        mask = crucible_mask(g) | (electrode_columns(g)[:, :, None] & (g.z > 0.6 * H_TOTAL)[None, None, :])
        rho = np.full((nr, nth, nz), 1.2)
        return assemble(g, ~mask, rho, _apu(rng, (nr, nth, nz)), rng, meta=meta)
    if config == 2:
        nr, nth, nz = shape or (60, 120, 84)
        g = make_geometry(nr, nth, nz)
        mask = crucible_mask(g) | (electrode_columns(g)[:, :, None] & (g.z > 0.6 * H_TOTAL)[None, None, :])
        rho = np.broadcast_to(np.where(g.z < 0.2 * H_TOTAL, 7000.0, 0.3)[None, None, :], (nr, nth, nz))
        return assemble(g, ~mask, rho, _apu(rng, (nr, nth, nz)), rng, meta=meta)
    if config == 3:
        if shape is None:
            nr = int(rng.integers(18, 36))
            nth = 6 * int(rng.integers(6, 12))       # 36..66, divisible by 6
            nz = int(rng.integers(24, 51))
        else:
            nr, nth, nz = shape
        h_liq = rng.uniform(0.05, 0.3) * H_TOTAL
        rho_g = rng.uniform(0.2, 1.5)
        g = make_geometry(nr, nth, nz)
        mask = crucible_mask(g) | (electrode_columns(g)[:, :, None] & (g.z > 0.6 * H_TOTAL)[None, None, :])
        rho = np.broadcast_to(np.where(g.z < h_liq, 7000.0, rho_g)[None, None, :], (nr, nth, nz))
        meta.update(h_liq=h_liq, rho_gas=rho_g)
        return assemble(g, ~mask, rho, _apu(rng, (nr, nth, nz)), rng, meta=meta)
    if config == 4:
        nr, nth, nz = shape or (120, 240, 168)
        g = make_geometry(nr, nth, nz)
        mask = crucible_mask(g) | (electrode_columns(g)[:, :, None] & (g.z > 0.6 * H_TOTAL)[None, None, :])
        rho = np.broadcast_to(np.where(g.z < 0.9, 7000.0, 0.3)[None, None, :], (nr, nth, nz))
        # scrap band 0.9 m < z < 2.5 m: Ergun factor 1e-2*(1 - alpha_s)
        alpha_s = rng.uniform(0.3, 0.6, size=(nr, nth, nz))
        band = ((g.z > 0.9) & (g.z < 2.5))[None, None, :]
        factor = np.where(band, 1e-2 * (1.0 - alpha_s), 1.0)
        meta.update(scrap_band=(0.9, 2.5))
        return assemble(g, ~mask, rho, _apu(rng, (nr, nth, nz)), rng, factor=factor, meta=meta)
    if config == 5:
        nr, nth, nz = shape or (60, 120, 84)
        g = make_geometry(nr, nth, nz)
        z_bath = 0.2 * H_TOTAL
        immersion = rng.uniform(0.2, 0.8)
        foam = rng.uniform(0.0, 0.4)
        lo = 10.0 ** (-3.0 + rng.uniform(-1.0, 1.0))
        hi = 10.0 ** (2.0 + rng.uniform(-1.0, 1.0))
        z_tip = z_bath - immersion
        mask = crucible_mask(g) | (electrode_columns(g)[:, :, None] & (g.z > z_tip)[None, None, :])
        rho_z = np.where(g.z < z_bath, 7000.0, np.where(g.z < z_bath + foam, 500.0, 0.3))
        rho = np.broadcast_to(rho_z[None, None, :], (nr, nth, nz))
        meta.update(immersion=immersion, foam=foam, apu_range=(lo, hi))
        return assemble(g, ~mask, rho, _apu(rng, (nr, nth, nz), lo, hi), rng, meta=meta)
    raise ValueError(f"unknown workload {config}")


def ensemble_shapes(config: int, count: int):
    """Mesh shapes of the first ``count`` systems of ensemble workload 3."""
    if config != 3:
        raise ValueError("only workload 3 has uneven meshes")
    out = []
    for s in range(count):
        rng = np.random.default_rng(system_seed(config, s))
        out.append((int(rng.integers(18, 36)), 6 * int(rng.integers(6, 12)), int(rng.integers(24, 51))))
    return out
