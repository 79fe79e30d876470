"""GPU parity of the mixed solver with the z-line inner preconditioner
(EAFP_PRECOND_ZLINE): the same bars as the Jacobi form (tests/
test_gpu_mixed.py), checked against the FP64 oracle of cg_3d_mpi on the
same systems:
  - true relative residual ||b - A x||/||b|| < 1e-8 with the caller's FP64
    operator (independent scipy operator on the host, 1.2e-8 for its
    rounding), agreeing with the solver's own FP64 residual;
  - relative p' error against the oracle (tol 1e-12) after the same gauge
    fix below 1e-5; zero mean on every singular component; inactive cells 0;
  - every tiling and stencil form gives a converged answer, a zero-RHS
    system stops at once and a NaN system never converges without holding
    the others back; a repeated solve repeats bit for bit;
  - on the production meshes it needs fewer inner iterations than Jacobi
    (the point of the preconditioner).
"""
import math

import numpy as np
import pytest

import eafp
import oracle
from eafp import systems
from eafp._lib import COEF_BF16, COEF_FP16, COEF_FP32, PRECOND_JACOBI, PRECOND_ZLINE, EAFP_ERR_UNSUPPORTED

pytestmark = pytest.mark.gpu

FORMATS = {"fp32": COEF_FP32, "bf16": COEF_BF16, "fp16": COEF_FP16}


def host_true_residual(ps, phi):
    A = oracle.csr(ps)
    x = oracle.interior(phi).ravel()
    b = ps.interior("Su").ravel()
    return np.linalg.norm(b - A @ x) / np.linalg.norm(b)


_ref = {}


def oracle_solution(ps, key):
    if key not in _ref:
        phi, _, nit, _ = oracle.cg(ps, max_iter=200000, tol=1e-12)
        _ref[key] = (oracle.interior(phi), nit)
    return _ref[key]


def check_answer(ps, phi, st, key):
    x_ref, _ = oracle_solution(ps, key)
    err = oracle.gauge_error(oracle.interior(phi), x_ref, ps.labels, ps.ncomp)
    tr = host_true_residual(ps, phi)
    print(f"  status={st['status']} inner={st['n_iter']} refine={st['n_refine']} res={st['residual']:.2e} "
          f"host_res={tr:.2e} err={err:.2e}", flush=True)
    assert st["status"] == "converged"
    assert st["residual"] < 1e-8 and tr < 1.2e-8
    assert abs(tr - st["residual"]) < 1e-3 * st["residual"] + 2e-12
    assert err < 1e-5
    x = oracle.interior(phi)
    for c in range(ps.ncomp):
        m = ps.labels == c
        assert abs(x[m].mean()) < 1e-12 * np.abs(x).max() + 1e-300
    assert not x[~ps.active].any()


@pytest.mark.parametrize("fmt", list(FORMATS))
@pytest.mark.parametrize("config,shape", [(1, None), (2, (20, 36, 28)), (3, (26, 54, 28)),
                                          (2, (20, 36, 29)), (3, (18, 36, 50)), (1, (12, 24, 18))])
def test_zline_single(config, shape, fmt):
    """nz a multiple of 4 (quad kernels) and not (generic kernels)."""
    ps = systems.make_system(config, 0, shape)
    print(f"case zline config={config} shape={ps.shape} fmt={fmt}", flush=True)
    with eafp.Batch.from_systems([ps]) as b:
        st = b.solve_mixed(tol=1e-8, coef_format=FORMATS[fmt], precond=PRECOND_ZLINE)[0]
        phi = b.download(0)
    check_answer(ps, phi, st, (config, shape))


@pytest.mark.parametrize("cpt,cols", [(2, 0), (8, 0), (4, 2)])
def test_zline_production_fewer_iterations(cpt, cols):
    """60x120x84 under every stencil form; fewer inner iterations than Jacobi."""
    ps = systems.make_system(2)
    print(f"case zline production cpt={cpt} columns={cols}", flush=True)
    with eafp.Batch.from_systems([ps]) as b:
        b.set_tiling(cpt, 0)
        b.set_columns(cols)
        stj = b.solve_mixed(tol=1e-8, coef_format=COEF_BF16, precond=PRECOND_JACOBI)[0]
        st = b.solve_mixed(tol=1e-8, coef_format=COEF_BF16, precond=PRECOND_ZLINE)[0]
        phi = b.download(0)
        stj2 = b.solve_mixed(tol=1e-8, coef_format=COEF_BF16, precond=PRECOND_JACOBI)[0]
    tr = host_true_residual(ps, phi)
    print(f"  jacobi inner={stj['n_iter']} | zline inner={st['n_iter']} refine={st['n_refine']} "
          f"res={st['residual']:.2e} host_res={tr:.2e}", flush=True)
    assert st["status"] == "converged" and tr < 1.2e-8
    assert st["n_iter"] < 0.9 * stj["n_iter"]
    # switching back restores the Jacobi run bit for bit
    assert stj2["n_iter"] == stj["n_iter"] and stj2["residual"] == stj["residual"]


def test_zline_batch_ragged_with_failures():
    """Uneven meshes (every nz a multiple of 4) in one launch sequence; a
    zero-RHS system returns at once and a NaN system never converges without
    stalling the others; a second solve repeats bit for bit."""
    syss = [systems.make_system(1, 0), systems.make_system(2, 0, (20, 36, 28)),
            systems.make_system(3, 0, (26, 54, 28)), systems.make_system(1, 1)]
    z = systems.make_system(1, 2)
    z.Su[...] = 0.0
    bad = systems.make_system(1, 3)
    bad.Su[6, 7, 8] = np.nan
    syss += [z, bad]
    with eafp.Batch.from_systems(syss) as b:
        st = b.solve_mixed(tol=1e-8, max_steps=100, coef_format=COEF_BF16, precond=PRECOND_ZLINE)
        for s, ps in enumerate(syss):
            print(f"case zline batch system={s} shape={ps.shape} -> {st[s]}", flush=True)
            phi = b.download(s)
            if s == 4:
                assert st[s]["status"] == "zero_rhs"
                continue
            if s == 5:
                assert st[s]["status"] == "max_iter" and math.isnan(st[s]["residual"])
                continue
            check_answer(ps, phi, st[s], ("zb", s))
        st2 = b.solve_mixed(tol=1e-8, max_steps=100, coef_format=COEF_BF16, precond=PRECOND_ZLINE)
        for a1, a2 in zip(st, st2):
            assert a1["n_iter"] == a2["n_iter"]
            assert (a1["residual"] == a2["residual"]) or (math.isnan(a1["residual"]) and math.isnan(a2["residual"]))


def test_zline_refused_when_ineligible():
    """nz above 1024: EAFP_ERR_UNSUPPORTED before any solve work."""
    ps = systems.make_system(3, 0)
    tall = systems.make_system(1, 0, (2, 4, 1028))
    with eafp.Batch.from_systems([ps, tall]) as b:
        with pytest.raises(eafp.EafpError) as e:
            b.solve_mixed(tol=1e-8, coef_format=COEF_BF16, precond=PRECOND_ZLINE)
        assert e.value.rc == EAFP_ERR_UNSUPPORTED
        # Jacobi still runs on the same batch (the 1028-cell columns converge slowly: only the other is checked)
        st = b.solve_mixed(tol=1e-8, coef_format=COEF_BF16, max_steps=200)
        assert st[0]["status"] == "converged"


def test_zline_pp_step():
    """The fused p' step with the z-line mixed solver meets the residual target."""
    from eafp import fields
    shape = (20, 36, 28)
    F, cfg = fields.make_fields(*shape, seed=1, masks=True)
    m = eafp.Mesh(*shape)
    out, pp, res, st = eafp.pp_step(F, cfg, m, 0.3, solver="mixed", mixed=dict(tol=1e-8, precond=PRECOND_ZLINE))
    outj, ppj, resj, stj = eafp.pp_step(F, cfg, m, 0.3, solver="mixed", mixed=dict(tol=1e-8))
    ref, _, _ = oracle.pp_assemble(F, cfg, shape)

    class S:
        pass
    ps = S()
    ps.shape = shape
    ps.padded = lambda n: ref[n]
    ps.interior = lambda n: oracle.interior(ref[n])
    tr = host_true_residual(ps, pp)
    print(f"case zline pp_step: inner {st['n_iter']} (jacobi {stj['n_iter']}) host_res={tr:.2e}", flush=True)
    assert st["status"] == "converged" and tr < 1.2e-8
    assert res == resj


def test_zline_ragged_any_nz():
    """The generator's ragged batch (nz 24..50, most not multiples of 4): the
    generic line kernels, fewer iterations than Jacobi on every system."""
    syss = [systems.make_system(3, s) for s in range(6)]
    assert any(p.shape[2] % 4 for p in syss)
    with eafp.Batch.from_systems(syss) as b:
        stj = b.solve_mixed(tol=1e-8, coef_format=COEF_BF16)
        st = b.solve_mixed(tol=1e-8, coef_format=COEF_BF16, precond=PRECOND_ZLINE)
        for s, ps in enumerate(syss):
            print(f"case zline ragged system={s} shape={ps.shape} jacobi={stj[s]['n_iter']} zline={st[s]['n_iter']}",
                  flush=True)
            check_answer(ps, b.download(s), st[s], ("rg", s))
            assert st[s]["n_iter"] < stj[s]["n_iter"]


def test_zline_autotune_and_kernel_timer():
    """Solver code 2 of the tuner and the kernel timer runs the z-line kernels;
    the batch solves correctly afterwards."""
    ps = systems.make_system(2, 0, (20, 36, 28))
    with eafp.Batch.from_systems([ps]) as b:
        tune = b.autotune(eafp.Batch.SOLVER_MIXED_ZLINE, COEF_BF16, reps=2)
        ms_sp, ms_up = b.time_kernels(eafp.Batch.SOLVER_MIXED_ZLINE, COEF_BF16, reps=3)
        print(f"case zline autotune {tune} spmv {ms_sp * 1e3:.1f} us update {ms_up * 1e3:.1f} us", flush=True)
        assert ms_sp > 0 and ms_up > 0
        st = b.solve_mixed(tol=1e-8, coef_format=COEF_BF16, precond=PRECOND_ZLINE)[0]
        phi = b.download(0)
    check_answer(ps, phi, st, (2, (20, 36, 28)))


def test_zline_invalid_precond_refused():
    """An unknown preconditioner code is EAFP_ERR_ARG; the batch still solves afterwards."""
    from eafp._lib import EAFP_ERR_ARG
    ps = systems.make_system(1, 0)
    with eafp.Batch.from_systems([ps]) as b:
        with pytest.raises(eafp.EafpError) as e:
            b.solve_mixed(tol=1e-8, coef_format=COEF_BF16, precond=7)
        assert e.value.rc == EAFP_ERR_ARG
        st = b.solve_mixed(tol=1e-8, coef_format=COEF_BF16, precond=PRECOND_ZLINE)[0]
        assert st["status"] == "converged"


def test_graph_survives_format_and_precond_switches():
    """Cached solver graphs name the buffers they were captured with.  A
    kernel timing in another weight format reallocates the weight buffers, and
    switching the preconditioner picks another graph: a BF16 z-line solve
    after FP32 timing, after a Jacobi solve and after an FP64 solve repeats the
    first BF16 z-line solve bit for bit, and the Jacobi solve repeats too."""
    ps = systems.make_system(2, 0, (20, 36, 28))
    with eafp.Batch.from_systems([ps]) as b:
        st0 = b.solve_mixed(tol=1e-8, coef_format=COEF_BF16, precond=PRECOND_ZLINE)[0]
        phi0 = b.download(0)
        b.time_kernels(eafp.Batch.SOLVER_MIXED, COEF_FP32, reps=2)
        st1 = b.solve_mixed(tol=1e-8, coef_format=COEF_BF16, precond=PRECOND_ZLINE)[0]
        phi1 = b.download(0)
        stj = b.solve_mixed(tol=1e-8, coef_format=COEF_BF16)[0]
        phij = b.download(0)
        b.time_kernels(eafp.Batch.SOLVER_MIXED_ZLINE, COEF_FP16, reps=2)
        b.solve_fp64(tol=1e-8)
        st2 = b.solve_mixed(tol=1e-8, coef_format=COEF_BF16, precond=PRECOND_ZLINE)[0]
        phi2 = b.download(0)
        stj2 = b.solve_mixed(tol=1e-8, coef_format=COEF_BF16)[0]
        phij2 = b.download(0)
    print(f"case graph switches: zline inner {st0['n_iter']} {st1['n_iter']} {st2['n_iter']} "
          f"jacobi {stj['n_iter']} {stj2['n_iter']}", flush=True)
    check_answer(ps, phi0, st0, (2, (20, 36, 28)))
    for st, phi in ((st1, phi1), (st2, phi2)):
        assert st["n_iter"] == st0["n_iter"] and st["residual"] == st0["residual"]
        np.testing.assert_array_equal(phi, phi0)
    assert stj2["n_iter"] == stj["n_iter"]
    np.testing.assert_array_equal(phij2, phij)
