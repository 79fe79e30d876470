"""GPU parity of the mixed-precision solver (FP32 inner CG on FP32/BF16/FP16
face weights, FP64 iterative refinement, FP64 zero-mean projection) against
the FP64 oracle of cg_3d_mpi on the same assembled systems.

Bar (the precision rule of the mixed path): true relative residual
||b - A x||/||b|| < 1e-8 with the caller's FP64 operator (checked on the
host with an independent scipy operator), the result zero-mean on every
singular component, and the relative p' error against an FP64 oracle run
(tol 1e-12) after the same gauge fix below 1e-5 (the error bound at equal
residual is the condition number times the residual; the measured values
are printed).
"""
import math

import numpy as np
import pytest

import eafp
import oracle
from eafp import systems
from eafp._lib import COEF_BF16, COEF_FP16, COEF_FP32

pytestmark = pytest.mark.gpu

FORMATS = {"fp32": COEF_FP32, "bf16": COEF_BF16, "fp16": COEF_FP16}


def host_true_residual(ps, phi):
    A = oracle.csr(ps)
    x = oracle.interior(phi).ravel()
    b = ps.interior("Su").ravel()
    return np.linalg.norm(b - A @ x) / np.linalg.norm(b)


_ref_cache = {}


def oracle_solution(ps, key):
    if key not in _ref_cache:
        phi, res, nit, _ = oracle.cg(ps, max_iter=200000, tol=1e-12)
        _ref_cache[key] = (oracle.interior(phi), nit)
    return _ref_cache[key]


@pytest.mark.parametrize("fmt", list(FORMATS))
@pytest.mark.parametrize("config,shape", [(1, None), (2, (20, 36, 28)), (3, None)])
def test_mixed_single(config, shape, fmt):
    ps = systems.make_system(config, 0, shape)
    print(f"case mixed config={config} shape={ps.shape} fmt={fmt}", flush=True)
    with eafp.Batch.from_systems([ps]) as b:
        st = b.solve_mixed(tol=1e-8, coef_format=FORMATS[fmt])[0]
        phi = b.download(0)
    x_ref, nit_ref = oracle_solution(ps, (config, shape))
    err = oracle.gauge_error(oracle.interior(phi), x_ref, ps.labels, ps.ncomp)
    tr = host_true_residual(ps, phi)
    print(f"  status={st['status']} inner={st['n_iter']} refine={st['n_refine']} "
          f"res={st['residual']:.2e} host_res={tr:.2e} l1={st['true_rel_l1']:.2e} err={err:.2e}", flush=True)
    assert st["status"] == "converged"
    assert st["residual"] < 1e-8
    assert tr < 1.2e-8
    assert abs(tr - st["residual"]) < 1e-3 * st["residual"] + 2e-12
    assert err < 1e-5
    x = oracle.interior(phi)
    assert abs(x[ps.active].mean()) < 1e-12 * np.abs(x).max() + 1e-300
    assert not x[~ps.active].any()


@pytest.mark.parametrize("cpt,cols", [(2, 0), (8, 0), (4, 2)])
def test_mixed_production_size(cpt, cols):
    """60x120x84 (nz a multiple of 4): cpt 2 runs the per-cell kernels, cpt 8
    the quad kernels, columns the column-marching stencil."""
    ps = systems.make_system(2)
    print(f"case mixed config=2 shape={ps.shape} cpt={cpt} columns={cols}", flush=True)
    with eafp.Batch.from_systems([ps]) as b:
        b.set_tiling(cpt, 0)
        b.set_columns(cols)
        st = b.solve_mixed(tol=1e-8, coef_format=COEF_BF16)[0]
        phi = b.download(0)
    tr = host_true_residual(ps, phi)
    print(f"  status={st['status']} inner={st['n_iter']} refine={st['n_refine']} res={st['residual']:.2e} "
          f"host_res={tr:.2e}", flush=True)
    assert st["status"] == "converged" and tr < 1.2e-8


def test_mixed_batch_ragged_with_failures():
    """Uneven meshes in one launch sequence; a zero-RHS system returns at once
    and a NaN system never converges without stalling the others."""
    syss = [systems.make_system(3, s) for s in range(5)]
    z = systems.make_system(1, 2)
    z.Su[...] = 0.0
    bad = systems.make_system(1, 3)
    bad.Su[6, 7, 8] = np.nan
    syss += [z, bad]
    with eafp.Batch.from_systems(syss) as b:
        st = b.solve_mixed(tol=1e-8, max_steps=100, coef_format=COEF_BF16)
        for s, ps in enumerate(syss):
            print(f"case mixed batch system={s} shape={ps.shape} -> {st[s]}", flush=True)
            phi = b.download(s)
            if s == 5:
                assert st[s]["status"] == "zero_rhs"
                continue
            if s == 6:
                assert st[s]["status"] == "max_iter" and math.isnan(st[s]["residual"])
                assert st[s]["n_refine"] == 100
                continue
            assert st[s]["n_iter"] <= 32 * st[s]["n_refine"]
            assert st[s]["status"] == "converged"
            assert host_true_residual(ps, phi) < 1.2e-8
            x_ref, _ = oracle_solution(ps, ("b3", s))
            assert oracle.gauge_error(oracle.interior(phi), x_ref, ps.labels, ps.ncomp) < 1e-5
        # deterministic: a second solve repeats bit for bit
        st2 = b.solve_mixed(tol=1e-8, max_steps=100, coef_format=COEF_BF16)
        for a1, a2 in zip(st, st2):
            assert a1["n_iter"] == a2["n_iter"]
            assert (a1["residual"] == a2["residual"]) or (math.isnan(a1["residual"]) and math.isnan(a2["residual"]))


def test_mixed_nonsingular_system_not_projected():
    """A reference-style regularised system (aP *= 1 + 1e-6, src/mod_pressure_3d.f90:261-265)
    is nonsingular: no component is projected and the answer matches the oracle."""
    ps = systems.make_system(1, 4)
    act = np.zeros_like(ps.coef["aP"], dtype=bool)
    act[2:-2, 2:-2, 2:-2] = ps.active
    ps.coef["aP"][act] *= 1.0 + 1e-6
    with eafp.Batch.from_systems([ps]) as b:
        st = b.solve_mixed(tol=1e-8, coef_format=COEF_FP32)[0]
        phi = b.download(0)
    phi_ref, _, _, _ = oracle.cg(ps, max_iter=200000, tol=1e-12)
    x, xr = oracle.interior(phi), oracle.interior(phi_ref)
    print(f"  {st} mean(x)={x[ps.active].mean():.3e}", flush=True)
    assert st["status"] == "converged"
    assert host_true_residual(ps, phi) < 1.2e-8
    print(f"  raw rel err {np.linalg.norm(x - xr) / np.linalg.norm(xr):.2e}", flush=True)
    assert oracle.gauge_error(x, xr, ps.labels, ps.ncomp) < 1e-4


@pytest.mark.parametrize("cpt", [1, 4, 8])
@pytest.mark.parametrize("restart", [False, True])
def test_mixed_tilings_and_restart(cpt, restart):
    syss = [systems.make_system(3, s) for s in range(3)]
    with eafp.Batch.from_systems(syss) as b:
        b.set_tiling(cpt, 0)
        st = b.solve_mixed(tol=1e-8, coef_format=COEF_FP16, restart=restart)
        for s, ps in enumerate(syss):
            phi = b.download(s)
            assert st[s]["status"] == "converged"
            assert host_true_residual(ps, phi) < 1.2e-8
            x_ref, _ = oracle_solution(ps, ("b3", s))
            assert oracle.gauge_error(oracle.interior(phi), x_ref, ps.labels, ps.ncomp) < 1e-5


@pytest.mark.parametrize("fmt", list(FORMATS))
@pytest.mark.parametrize("cpt", [4, 8])
def test_mixed_quad_kernels(cpt, fmt):
    """Batches whose every nz is a multiple of 4 run the quad kernels at cpt 4
    and 8 (4 cells of a z-row per thread): same bar as the per-cell kernels,
    over ragged shapes with partial last tiles, nz = 4 (one quad per row, both
    z walls in one quad) and a single theta cell (the wrap is the cell row
    itself)."""
    syss = [systems.make_system(2, 0, (20, 36, 28)), systems.make_system(1, 0),
            systems.make_system(5, 1, (24, 48, 32)), systems.make_system(2, 1, (6, 10, 4)),
            systems.make_system(1, 5, (5, 1, 12))]
    with eafp.Batch.from_systems(syss) as b:
        b.set_tiling(cpt, 0)
        st = b.solve_mixed(tol=1e-8, coef_format=FORMATS[fmt])
        for s, ps in enumerate(syss):
            phi = b.download(s)
            tr = host_true_residual(ps, phi)
            x_ref, _ = oracle_solution(ps, ("quad", s))
            err = oracle.gauge_error(oracle.interior(phi), x_ref, ps.labels, ps.ncomp)
            print(f"case quad cpt={cpt} fmt={fmt} system={s} shape={ps.shape}: {st[s]['status']} "
                  f"inner={st[s]['n_iter']} res={tr:.2e} err={err:.2e}", flush=True)
            assert st[s]["status"] == "converged"
            assert tr < 1.2e-8
            assert err < 1e-5
            x = oracle.interior(phi)
            if ps.ncomp:
                assert abs(x[ps.active].mean()) < 1e-12 * np.abs(x).max() + 1e-300


@pytest.mark.parametrize("fmt", list(FORMATS))
@pytest.mark.parametrize("splits", [1, 2, 3])
def test_mixed_column_kernel(splits, fmt):
    """The column-marching stencil (theta-row slabs marched through r, one
    quad per thread, LDS theta/z neighbours): same bar as the tile kernels.
    Shapes: a slab narrower than nth with a partial last column, a single
    column whose halo rows wrap onto itself (nth*nz <= 1024), nz = 32 and 28,
    r-segments that split the planes unevenly.  Odd segments march down in
    r: with 2 splits one ends at the outer wall, with 3 one lies between two
    upward ones."""
    syss = [systems.make_system(2, 0, (20, 36, 28)), systems.make_system(1, 0),
            systems.make_system(5, 1, (24, 48, 32)), systems.make_system(2, 3, (7, 50, 20))]
    with eafp.Batch.from_systems(syss) as b:
        b.set_tiling(1, 0)
        b.set_columns(splits)
        assert b.info()["column_splits"] == splits
        st = b.solve_mixed(tol=1e-8, coef_format=FORMATS[fmt])
        for s, ps in enumerate(syss):
            phi = b.download(s)
            tr = host_true_residual(ps, phi)
            x_ref, _ = oracle_solution(ps, ("col", s))
            err = oracle.gauge_error(oracle.interior(phi), x_ref, ps.labels, ps.ncomp)
            print(f"case columns splits={splits} fmt={fmt} system={s} shape={ps.shape}: {st[s]['status']} "
                  f"inner={st[s]['n_iter']} res={tr:.2e} err={err:.2e}", flush=True)
            assert st[s]["status"] == "converged"
            assert tr < 1.2e-8
            assert err < 1e-5
            x = oracle.interior(phi)
            if ps.ncomp:
                assert abs(x[ps.active].mean()) < 1e-12 * np.abs(x).max() + 1e-300


def test_mixed_column_kernel_ragged_with_failures():
    """Column path with a zero-RHS and a NaN system in the batch: both end as
    in the tile path, the others converge."""
    syss = [systems.make_system(2, 0, (20, 36, 28)), systems.make_system(1, 1)]
    z = systems.make_system(1, 2)
    z.Su[...] = 0.0
    bad = systems.make_system(1, 3)
    bad.Su[6, 7, 8] = np.nan
    syss += [z, bad]
    with eafp.Batch.from_systems(syss) as b:
        b.set_columns(1)
        st = b.solve_mixed(tol=1e-8, max_steps=100, coef_format=COEF_FP32)
        print(f"case columns ragged: {[x['status'] for x in st]}", flush=True)
        assert st[2]["status"] == "zero_rhs"
        assert st[3]["status"] == "max_iter" and math.isnan(st[3]["residual"])
        for s in (0, 1):
            assert st[s]["status"] == "converged"
            assert host_true_residual(syss[s], b.download(s)) < 1.2e-8


def test_mixed_columns_refused_when_ineligible():
    """nz above 256: the column setting is refused and the tile kernels stay
    in use."""
    ps = systems.make_system(2, 0, (4, 6, 264))
    with eafp.Batch.from_systems([ps]) as b:
        with pytest.raises(eafp.EafpError):
            b.set_columns(1)
        assert b.info()["column_splits"] == 0


@pytest.mark.parametrize("fmt", list(FORMATS))
@pytest.mark.parametrize("splits", [1, 4])
def test_mixed_generic_column_kernel(splits, fmt):
    """A batch with nz not a multiple of 4 runs the generic column kernel
    (4 strided cells per thread, z neighbours from LDS): ragged workload-3
    shapes, a zero-RHS and a NaN system."""
    syss = [systems.make_system(3, s) for s in range(4)]
    z = systems.make_system(2, 2, (6, 12, 13))
    z.Su[...] = 0.0
    bad = systems.make_system(2, 3, (6, 12, 13))
    bad.Su[4, 5, 6] = np.nan
    syss += [z, bad]
    assert any(ps.shape[2] % 4 for ps in syss)
    with eafp.Batch.from_systems(syss) as b:
        b.set_tiling(1, 0)
        b.set_columns(splits)
        st = b.solve_mixed(tol=1e-8, max_steps=200, coef_format=FORMATS[fmt])
        assert st[4]["status"] == "zero_rhs"
        assert st[5]["status"] == "max_iter" and math.isnan(st[5]["residual"])
        for s in range(4):
            ps = syss[s]
            phi = b.download(s)
            tr = host_true_residual(ps, phi)
            x_ref, _ = oracle_solution(ps, ("b3", s))
            err = oracle.gauge_error(oracle.interior(phi), x_ref, ps.labels, ps.ncomp)
            print(f"case generic columns splits={splits} fmt={fmt} system={s} shape={ps.shape}: "
                  f"{st[s]['status']} inner={st[s]['n_iter']} res={tr:.2e} err={err:.2e}", flush=True)
            assert st[s]["status"] == "converged"
            assert tr < 1.2e-8
            assert err < 1e-5
