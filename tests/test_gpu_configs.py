"""GPU parity of the mixed solver on the two benchmark workloads the other
tests only run scaled down: the 120x240x168 system with a scrap band
(coefficient contrast about 1e6, BF16 weights) and the 60x120x84
campaign-variant ensemble (electrode immersion, slag foam, jittered a_P^u),
held to the FP64 oracle of cg_3d_mpi (src/mod_solver_3d.f90:485-620):

  - scaled 120x240x168 (30x60x42, the same scrap band and pool): relative p'
    error against the oracle run to 1e-12 below 1e-5, true relative residual
    below 1e-8 checked on the host with scipy's CSR operator;
  - full 120x240x168: the true relative residual with scipy's CSR operator,
    and the p' error against the GPU FP64 Jacobi-PCG run to 1e-12 (the
    drop-in's algorithm, itself held to the oracle by test_gpu_cg64.py);
  - two full-size 60x120x84 ensemble systems in one batch, both inner
    preconditioners: p' error against the multi-threaded oracle run to
    1e-12, host true residual, zero mean on every singular component.
"""
import numpy as np
import pytest

import eafp
import oracle
from eafp import systems
from eafp._lib import COEF_BF16, PRECOND_JACOBI, PRECOND_ZLINE

pytestmark = pytest.mark.gpu


def host_true_residual(ps, phi):
    A = oracle.csr(ps)
    x = oracle.interior(phi).ravel()
    b = ps.interior("Su").ravel()
    return float(np.linalg.norm(b - A @ x) / np.linalg.norm(b))


def check(ps, phi, st, x_ref, label):
    err = oracle.gauge_error(oracle.interior(phi), x_ref, ps.labels, ps.ncomp)
    tr = host_true_residual(ps, phi)
    print(f"  {label}: status={st['status']} inner={st['n_iter']} refine={st['n_refine']} res={st['residual']:.2e} "
          f"host_res={tr:.2e} l1={st['true_rel_l1']:.2e} err={err:.2e}", flush=True)
    assert st["status"] == "converged"
    assert st["residual"] < 1e-8 and tr < 1.2e-8
    assert err < 1e-5
    x = oracle.interior(phi)
    for c in range(ps.ncomp):
        m = ps.labels == c
        assert abs(x[m].mean()) <= 1e-12 * np.abs(x[m]).max()
    assert not x[~ps.active].any()


@pytest.mark.parametrize("precond", [PRECOND_ZLINE, PRECOND_JACOBI])
def test_scrap_band_scaled_vs_oracle(precond):
    ps = systems.make_system(4, 0, (30, 60, 42))
    print(f"case config4 scaled shape={ps.shape} precond={precond}", flush=True)
    phi_ref, res_ref, nit_ref, _ = oracle.cg(ps, max_iter=400000, tol=1e-12)
    assert res_ref < 1e-12
    with eafp.Batch.from_systems([ps]) as b:
        st = b.solve_mixed(tol=1e-8, coef_format=COEF_BF16, precond=precond)[0]
        phi = b.download(0)
    check(ps, phi, st, oracle.interior(phi_ref), f"oracle n_iter={nit_ref}")


def test_scrap_band_full_size():
    ps = systems.make_system(4, 0)
    print(f"case config4 full shape={ps.shape}", flush=True)
    with eafp.Batch.from_systems([ps]) as b:
        b.set_columns(12)
        st = b.solve_mixed(tol=1e-8, coef_format=COEF_BF16, precond=PRECOND_ZLINE)[0]
        phi = b.download(0)
        st64 = b.solve_fp64(400000, 1e-12)[0]
        phi64 = b.download(0)
    assert st64["status"] == "converged"
    check(ps, phi, st, oracle.interior(phi64), f"fp64 1e-12 n_iter={st64['n_iter']}")


def test_campaign_ensemble_full_size_vs_oracle():
    syss = [systems.make_system(5, s) for s in (0, 1)]
    refs = []
    for ps in syss:
        phi_ref, res_ref, nit_ref = oracle.cg_omp(ps, max_iter=400000, tol=1e-12)
        assert res_ref < 1e-12
        refs.append(oracle.interior(phi_ref))
    with eafp.Batch.from_systems(syss) as b:
        for precond in (PRECOND_ZLINE, PRECOND_JACOBI):
            st = b.solve_mixed(tol=1e-8, coef_format=COEF_BF16, precond=precond)
            for s, ps in enumerate(syss):
                print(f"case config5 full system={s} shape={ps.shape} precond={precond} "
                      f"immersion={ps.meta['immersion']:.2f} foam={ps.meta['foam']:.2f}", flush=True)
                check(ps, b.download(s), st[s], refs[s], "oracle 1e-12")
