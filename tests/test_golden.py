"""Regression fixtures of tests/golden/ (written by tests/golden/make_golden.py
with this repository's own oracle; the reference has no golden vectors for its
p' CG): the generator must reproduce the inputs bit for bit, the oracle its
own result, and the GPU drop-in the oracle's result to the parity bar."""
import hashlib
import json
import os

import numpy as np
import pytest

import eafp
import oracle
from eafp import systems

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = ["config1_12x24x16", "config2_20x36x28"]


def load(name):
    with open(os.path.join(HERE, f"{name}.json")) as f:
        meta = json.load(f)
    phi = np.load(os.path.join(HERE, f"{name}_phi.npy"), allow_pickle=False)
    shape = tuple(meta["shape"]) if meta["config"] != 1 else None
    ps = systems.make_system(meta["config"], meta["index"], shape)
    return meta, phi, ps


@pytest.mark.parametrize("name", CASES)
def test_generator_reproduces_golden_inputs(name):
    meta, _, ps = load(name)
    for n, h in meta["sha256"].items():
        assert hashlib.sha256(np.asfortranarray(ps.padded(n)).tobytes(order="F")).hexdigest() == h, n


@pytest.mark.parametrize("name", CASES)
def test_oracle_reproduces_golden_result(name):
    meta, phi_ref, ps = load(name)
    phi, res, nit, hist = oracle.cg(ps, max_iter=meta["max_iter"], tol=meta["tol"], history=True)
    assert nit == meta["n_iter"]
    np.testing.assert_allclose(hist[:len(meta["residual_history_head"])], meta["residual_history_head"],
                               rtol=1e-12, atol=0)
    x = oracle.interior(phi)
    assert np.linalg.norm(x - phi_ref) <= 1e-10 * np.linalg.norm(phi_ref)


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_gpu_dropin_matches_golden(name):
    meta, phi_ref, ps = load(name)
    nr, nth, nz = ps.shape
    phi = np.zeros((nr + 4, nth + 4, nz + 4), order="F")
    arrs = [ps.padded(n) for n in ("aW", "aE", "aS", "aN", "aB", "aT", "aP", "Su")]
    res, nit = eafp.cg_3d_mpi(*arrs, phi, eafp.Mesh(nr, nth, nz), meta["max_iter"], meta["tol"])
    print(f"case golden {name}: gpu n_iter={nit} res={res:.3e} | golden n_iter={meta['n_iter']}", flush=True)
    assert res < meta["tol"]
    assert abs(nit - meta["n_iter"]) <= max(2, 0.02 * meta["n_iter"])
    err = oracle.gauge_error(oracle.interior(phi), phi_ref, ps.labels, ps.ncomp)
    assert err < 1e-6


# ---------------------------------------------------------------------------
# bit-exact paths: p' assembly with face exports, post-solve steps, line TDMA

PP_CASES = ["pp_12x24x16", "pp_16x30x24_single_phase"]


def sha(a):
    return hashlib.sha256(np.asfortranarray(a).tobytes(order="F")).hexdigest()


def load_pp(name):
    from eafp import fields
    with open(os.path.join(HERE, f"{name}.json")) as f:
        meta = json.load(f)
    shape = tuple(meta["shape"])
    F, cfg = fields.make_fields(*shape, seed=meta["seed"], masks=True, two_phase=meta["two_phase"])
    rng = np.random.default_rng(900 + meta["seed"])
    pp = fields.theta_halo(np.asfortranarray(50.0 * rng.normal(size=F["p"].shape)))
    return meta, shape, F, cfg, pp


@pytest.mark.parametrize("name", PP_CASES)
def test_pp_golden_inputs_and_oracle(name):
    meta, shape, F, cfg, pp = load_pp(name)
    assert {k: sha(v) for k, v in sorted(F.items())} == meta["inputs"]
    assert sha(pp) == meta["pp"] and cfg == meta["config"]
    coef, res, flux, ex = oracle.pp_assemble(F, cfg, shape, exports=True)
    assert {k: sha(v) for k, v in coef.items()} == meta["assembly"]
    assert {k: sha(v) for k, v in ex.items()} == meta["exports"]
    assert res == meta["residual"] and flux == meta["flux_ref"]
    upd = oracle.pp_correct(F, cfg, shape, meta["alpha_p"], pp, ex)
    assert {k: sha(v) for k, v in upd.items()} == meta["post_solve"]
    assert sha(oracle.tdma_3d(coef, np.zeros_like(coef["aP"], order="F"), shape, 2)) == meta["tdma_2_sweeps"]


@pytest.mark.gpu
@pytest.mark.parametrize("name", PP_CASES)
def test_pp_golden_gpu(name):
    """The GPU reproduces every digest of the fixture: bit-identical results."""
    meta, shape, F, cfg, pp = load_pp(name)
    m = eafp.Mesh(*shape)
    coef, res, flux, ex = eafp.pp_assemble(F, cfg, m, exports=True)
    print(f"case golden {name}", flush=True)
    assert {k: sha(v) for k, v in coef.items()} == meta["assembly"]
    assert {k: sha(v) for k, v in ex.items()} == meta["exports"]
    assert abs(res - meta["residual"]) <= 1e-12 * abs(meta["residual"])
    upd = eafp.pp_correct(F, cfg, m, meta["alpha_p"], pp, ex)
    assert {k: sha(v) for k, v in upd.items()} == meta["post_solve"]
    phi = np.zeros_like(coef["aP"], order="F")
    eafp.tdma_3d_mpi(*[coef[n] for n in ("aW", "aE", "aS", "aN", "aB", "aT", "aP", "Su")], phi, m, 2)
    assert sha(phi) == meta["tdma_2_sweeps"]
