"""GPU parity of the line TDMA (eafp_tdma_solve, eafp_tdma_solve_batch)
against the oracle's restatement of tdma_3d_mpi (oracle/tdma.c, pinned by
tests/test_tdma_oracle.py).

Bar: phi bit-identical to the oracle on every padded cell, halos included
(the GPU keeps the reference's line order exactly and evaluates the same
expressions without contraction), after 1 to 5 sweeps, on random diagonally
dominant systems, rows with zero pivots, the production p' system, ragged
batches, degenerate axes and device-resident arrays; through both kernels
(k_tdma_chain where it plans a pass, k_tdma_pass for lines of one element,
pivots outside the chain kernel's band, and EAFP_TDMA_KERNEL=pass).
"""
import ctypes

import numpy as np
import pytest

import eafp
import oracle
from eafp import _lib, systems
from hipbuf import HipBuffers

pytestmark = pytest.mark.gpu

NAMES = ("aW", "aE", "aS", "aN", "aB", "aT", "aP", "Su")


def random_system(shape, seed, zero_pivots=False, huge=False):
    nr, nth, nz = shape
    rng = np.random.default_rng(seed)
    pad = (nr + 4, nth + 4, nz + 4)
    coef = {n: np.asfortranarray(rng.uniform(0.1, 1.0, pad)) for n in NAMES[:6]}
    coef["aP"] = np.asfortranarray(sum(coef[n] for n in NAMES[:6]) * (1.0 + rng.uniform(0.0, 0.3, pad)))
    coef["Su"] = np.asfortranarray(rng.normal(size=pad))
    if zero_pivots:
        flat = [coef[n].reshape(-1, order="F") for n in NAMES]
        sel = np.arange(0, flat[0].size, 7)
        for a in flat[:6]:
            a[sel] = 0.0
        flat[6][sel] = np.where(sel % 2 == 1, 0.0, -0.0)
    if huge:
        # pivots near the top of the double range: the division's operands get rescaled (the GPU's
        # fallback path); tiny right-hand sides give quotients near the bottom
        flat = coef["aP"].reshape(-1, order="F")
        sel = rng.random(flat.size) < 0.02
        flat[sel] = 4.0e307 * rng.uniform(1.0, 2.0, np.count_nonzero(sel))
        su = coef["Su"].reshape(-1, order="F")
        tiny = rng.random(su.size) < 0.02
        su[tiny] = 1e-300 * rng.normal(size=np.count_nonzero(tiny))
    phi = np.asfortranarray(rng.normal(size=pad))
    return coef, phi


def assert_bitwise(got, want, what):
    bad = np.count_nonzero(got.view(np.uint64) != want.view(np.uint64))
    assert bad == 0, f"{what}: {bad} of {got.size} padded values differ, max |diff| {np.abs(got - want).max():.3e}"


CASES = [((12, 24, 16), 1, False), ((12, 24, 16), 3, True), ((20, 36, 28), 2, False), ((7, 12, 9), 5, True),
         ((1, 5, 4), 2, False), ((6, 1, 3), 2, True), ((4, 6, 1), 2, False), ((3, 2, 5), 3, False),
         ((1, 1, 1), 2, False), ((60, 120, 84), 1, False), ((300, 6, 280), 1, False)]


@pytest.mark.parametrize("shape,n_sweep,zero_pivots", CASES)
def test_tdma_matches_oracle_bitwise(shape, n_sweep, zero_pivots):
    print(f"case tdma shape={shape} n_sweep={n_sweep} zero_pivots={zero_pivots}", flush=True)
    coef, phi0 = random_system(shape, sum(shape) + n_sweep, zero_pivots)
    want = oracle.tdma_3d(coef, phi0, shape, n_sweep)
    phi = phi0.copy(order="F")
    eafp.tdma_3d_mpi(*[coef[n] for n in NAMES], phi, eafp.Mesh(*shape), n_sweep)
    assert_bitwise(phi, want, f"{shape} x{n_sweep}")


@pytest.mark.parametrize("shape", [(12, 24, 16), (30, 20, 25)])
def test_tdma_extreme_magnitudes(shape):
    """Pivots near 1e308 and tiny sources: divisions whose operands the hardware rescales."""
    print(f"case tdma extreme shape={shape}", flush=True)
    coef, phi0 = random_system(shape, 77 + shape[0], huge=True)
    want = oracle.tdma_3d(coef, phi0, shape, 2)
    phi = phi0.copy(order="F")
    eafp.tdma_3d_mpi(*[coef[n] for n in NAMES], phi, eafp.Mesh(*shape), 2)
    assert_bitwise(phi, want, f"extreme {shape}")


def test_tdma_production_system():
    """The 60x120x84 p' system (walls, outlet rows, liquid/gas jump) from phi = 0."""
    ps = systems.make_system(2)
    shape = ps.shape
    coef = {n: ps.padded(n) for n in NAMES}
    phi0 = np.zeros_like(coef["aP"], order="F")
    want = oracle.tdma_3d(coef, phi0, shape, 2)
    phi = phi0.copy(order="F")
    eafp.tdma_3d_mpi(*[coef[n] for n in NAMES], phi, eafp.Mesh(*shape), 2)
    print(f"case tdma production: max|phi| {np.abs(phi).max():.3e}", flush=True)
    assert_bitwise(phi, want, "production")


def test_tdma_batch_ragged():
    shapes = [(12, 24, 16), (5, 9, 30), (1, 3, 2), (17, 40, 8), (9, 1, 11), (33, 20, 25)]
    sys_, meshes, wants = [], [], []
    for q, shape in enumerate(shapes):
        coef, phi0 = random_system(shape, 100 + q, zero_pivots=(q % 2 == 1))
        wants.append(oracle.tdma_3d(coef, phi0, shape, 3))
        d = dict(coef)
        d["phi"] = phi0.copy(order="F")
        sys_.append(d)
        meshes.append(eafp.Mesh(*shape))
    print(f"case tdma batch of {len(shapes)}", flush=True)
    eafp.tdma_batch(sys_, meshes, 3)
    for q in range(len(shapes)):
        assert_bitwise(sys_[q]["phi"], wants[q], f"system {q} {shapes[q]}")


def test_tdma_batch_ragged_chain_entries():
    """Every pass of every system runs the chain kernel, whose helper lanes
    take as many ring entries per chunk as the batch's largest plan needs
    (1 to 3 here), and whose pack tiles are cut at every mesh edge."""
    shapes = [(12, 24, 16), (60, 30, 20), (17, 40, 8), (33, 20, 25), (2, 70, 3)]
    sys_, meshes, wants = [], [], []
    for q, shape in enumerate(shapes):
        coef, phi0 = random_system(shape, 300 + q)
        wants.append(oracle.tdma_3d(coef, phi0, shape, 2))
        d = dict(coef)
        d["phi"] = phi0.copy(order="F")
        sys_.append(d)
        meshes.append(eafp.Mesh(*shape))
    print(f"case tdma chain batch of {len(shapes)}", flush=True)
    eafp.tdma_batch(sys_, meshes, 2)
    for q in range(len(shapes)):
        assert_bitwise(sys_[q]["phi"], wants[q], f"system {q} {shapes[q]}")


def test_tdma_device_resident_and_repeat():
    """arrays_on_device = 1; two calls in a row continue from the first one's phi."""
    shape = (20, 36, 28)
    coef, phi0 = random_system(shape, 7)
    want = oracle.tdma_3d(coef, oracle.tdma_3d(coef, phi0, shape, 2), shape, 1)
    pad = tuple(n + 4 for n in shape)
    hip = HipBuffers()
    try:
        ptrs = [hip.put(coef[n]) for n in NAMES]
        phi_d = hip.put(phi0)
        L = _lib.load()
        d = eafp.solver.desc_of(eafp.Mesh(*shape))
        _lib.check(L.eafp_tdma_solve(ctypes.byref(d), *ptrs, phi_d, 2, 1), "eafp_tdma_solve")
        _lib.check(L.eafp_tdma_solve(ctypes.byref(d), *ptrs, phi_d, 1, 1), "eafp_tdma_solve")
        assert_bitwise(hip.get(phi_d, pad, np.float64), want, "device-resident")
    finally:
        hip.free()


@pytest.mark.parametrize("shape", [(12, 24, 16), (60, 120, 84)])
def test_tdma_both_kernels(shape, monkeypatch):
    """The same call through k_tdma_pass (EAFP_TDMA_KERNEL=pass) and through
    the chain kernel (the default wherever it plans the pass): both
    bit-identical to the oracle."""
    coef, phi0 = random_system(shape, 300 + shape[0])
    want = oracle.tdma_3d(coef, phi0, shape, 2)
    for kernel in ("pass", ""):
        print(f"case tdma kernel={kernel or 'chain'} shape={shape}", flush=True)
        monkeypatch.setenv("EAFP_TDMA_KERNEL", kernel)
        phi = phi0.copy(order="F")
        eafp.tdma_3d_mpi(*[coef[n] for n in NAMES], phi, eafp.Mesh(*shape), 2)
        assert_bitwise(phi, want, f"{kernel or 'chain'} {shape}")
