"""GPU parity of the p' assembler (eafp_pp_assemble) against the oracle's
restatement (oracle/pp_assemble.c, pinned by tests/test_pp_oracle.py).

Bar: every coefficient array and Su bit-identical (integer-like work on
doubles: the same operations in the same order, no contraction); the
continuity residual and flux_ref, whose global sums are reduced in another
order, within 1e-12 relative.  End to end, the GPU-assembled system solved
by the GPU drop-in CG matches the oracle CG on the oracle-assembled system.
"""
import numpy as np
import pytest

import eafp
import oracle
from eafp import fields

pytestmark = pytest.mark.gpu

NAMES = ("aW", "aE", "aS", "aN", "aB", "aT", "aP", "Su")
CASES = [((12, 24, 16), dict(seed=0)),
         ((20, 36, 28), dict(seed=1, masks=True)),
         ((16, 30, 24), dict(seed=2, two_phase=False)),
         ((24, 48, 32), dict(seed=3, scrap=False)),
         ((7, 12, 9), dict(seed=4, masks=True)),
         ((60, 120, 84), dict(seed=5, masks=True))]


@pytest.mark.parametrize("shape,kw", CASES)
def test_assembly_matches_oracle_bitwise(shape, kw):
    print(f"case assemble shape={shape} {kw}", flush=True)
    F, cfg = fields.make_fields(*shape, **kw)
    ref, res_ref, flux_ref = oracle.pp_assemble(F, cfg, shape)
    out, res, flux = eafp.pp_assemble(F, cfg, eafp.Mesh(*shape))
    for n in NAMES:
        bad = np.count_nonzero(out[n] != ref[n])
        assert bad == 0, f"{n}: {bad} cells differ, max |diff| {np.abs(out[n] - ref[n]).max():.3e}"
    assert abs(flux - flux_ref) <= 1e-12 * flux_ref
    assert abs(res - res_ref) <= 1e-12 * max(res_ref, 1e-300)
    print(f"  residual {res:.6e} (oracle {res_ref:.6e}) flux_ref {flux:.6e}", flush=True)


@pytest.mark.parametrize("flags", [dict(gas_in_poisson=0), dict(gas_compressibility=0), dict(solve_multiphase=0)])
def test_assembly_flags(flags):
    shape = (14, 24, 20)
    F, cfg = fields.make_fields(*shape, seed=6, masks=True)
    cfg = dict(cfg, **flags)
    ref, res_ref, _ = oracle.pp_assemble(F, cfg, shape)
    out, res, _ = eafp.pp_assemble(F, cfg, eafp.Mesh(*shape))
    for n in NAMES:
        np.testing.assert_array_equal(out[n], ref[n], err_msg=n)
    assert abs(res - res_ref) <= 1e-12 * max(res_ref, 1e-300)


def test_assemble_then_solve_matches_oracle_pipeline():
    """GPU assembly -> GPU drop-in CG against oracle assembly -> oracle CG."""
    shape = (20, 36, 28)
    F, cfg = fields.make_fields(*shape, seed=7, masks=True)
    m = eafp.Mesh(*shape)
    out, res, _ = eafp.pp_assemble(F, cfg, m)
    phi = np.zeros_like(out["aP"], order="F")
    r_gpu, n_gpu = eafp.cg_3d_mpi(*[out[n] for n in NAMES], phi, m, 20000, 1e-10)
    ref, _, _ = oracle.pp_assemble(F, cfg, shape)

    class S:
        pass
    ps = S()
    ps.shape = shape
    ps.padded = lambda n: ref[n]
    phi_ref, r_ref, n_ref, _ = oracle.cg(ps, max_iter=20000, tol=1e-10)
    print(f"case pipeline: gpu n_iter={n_gpu} res={r_gpu:.2e} | oracle n_iter={n_ref} res={r_ref:.2e}", flush=True)
    assert r_gpu < 1e-10 and abs(n_gpu - n_ref) <= max(2, 0.02 * n_ref)
    x, xr = oracle.interior(phi), oracle.interior(phi_ref)
    assert np.linalg.norm(x - xr) <= 1e-6 * np.linalg.norm(xr)
