"""Writes the regression fixtures of tests/golden/ (run from the repository root:
``python tests/golden/make_golden.py``).

The reference holds no golden vectors or known-answer tests for its p' CG
(cg_3d_mpi, src/mod_solver_3d.f90:485-620), its p' assembly and post-solve
steps, or its line TDMA beyond tests/unit/test_tdma.f90 (restated in
tests/test_tdma_oracle.py), and its Fortran/MPI build cannot be made in this
image, so these fixtures are written by this repository's own FP64 oracle
(oracle/).  They pin two things against silent
change: the seeded system generator (a SHA-256 of every padded input array)
and the oracle's result on those inputs (iteration count, residual, the first
residuals of the history and the whole phi interior).  The oracle itself is
pinned independently by tests/test_oracle.py.
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "hornofusion-full3d_amd"))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from eafp import systems  # noqa: E402

CASES = [("config1_12x24x16", 1, 0, None), ("config2_20x36x28", 2, 0, (20, 36, 28))]
ARRAYS = ("aW", "aE", "aS", "aN", "aB", "aT", "aP", "Su")
TOL = 1e-8
MAX_ITER = 100000


def digest(ps):
    return {n: hashlib.sha256(np.asfortranarray(ps.padded(n)).tobytes(order="F")).hexdigest() for n in ARRAYS}


PP_CASES = [("pp_12x24x16", (12, 24, 16), 0, True), ("pp_16x30x24_single_phase", (16, 30, 24), 2, False)]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes(order="F") if a.flags["F_CONTIGUOUS"]
                          else np.asfortranarray(a).tobytes(order="F")).hexdigest()


def pp_inputs(shape, seed, two_phase):
    """The seeded fields, the correction field p' (theta halos periodic) and
    alpha_p of one assembly / post-solve / line-TDMA fixture."""
    from eafp import fields
    F, cfg = fields.make_fields(*shape, seed=seed, masks=True, two_phase=two_phase)
    rng = np.random.default_rng(900 + seed)
    pp = fields.theta_halo(np.asfortranarray(50.0 * rng.normal(size=F["p"].shape)))
    return F, cfg, pp, 0.3


def pp_main():
    """Bit-exact paths: the fixture holds SHA-256 digests of the inputs and of
    the oracle's outputs (assembled coefficients and face exports, post-solve
    fluxes, links, velocities and p, and phi after two line-TDMA sweeps of the
    assembled system from zero), which the GPU must reproduce bit for bit."""
    for name, shape, seed, two_phase in PP_CASES:
        F, cfg, pp, alpha_p = pp_inputs(shape, seed, two_phase)
        coef, res, flux, ex = oracle.pp_assemble(F, cfg, shape, exports=True)
        upd = oracle.pp_correct(F, cfg, shape, alpha_p, pp, ex)
        tdma = oracle.tdma_3d(coef, np.zeros_like(coef["aP"], order="F"), shape, 2)
        meta = {"shape": list(shape), "seed": seed, "two_phase": two_phase, "alpha_p": alpha_p, "config": cfg,
                "inputs": {k: sha(v) for k, v in sorted(F.items())}, "pp": sha(pp),
                "assembly": {k: sha(v) for k, v in coef.items()}, "exports": {k: sha(v) for k, v in ex.items()},
                "residual": res, "flux_ref": flux, "post_solve": {k: sha(v) for k, v in upd.items()},
                "tdma_2_sweeps": sha(tdma)}
        with open(os.path.join(HERE, f"{name}.json"), "w") as f:
            json.dump(meta, f, indent=1)
            f.write("\n")
        print(f"{name}: residual={res:.6e}")


def main():
    for name, config, index, shape in CASES:
        ps = systems.make_system(config, index, shape)
        phi, res, nit, hist = oracle.cg(ps, max_iter=MAX_ITER, tol=TOL, history=True)
        meta = {"config": config, "index": index, "shape": list(ps.shape), "tol": TOL, "max_iter": MAX_ITER,
                "sha256": digest(ps), "n_iter": nit, "residual": res,
                "residual_history_head": [float(v) for v in hist[:20]]}
        with open(os.path.join(HERE, f"{name}.json"), "w") as f:
            json.dump(meta, f, indent=1)
            f.write("\n")
        np.save(os.path.join(HERE, f"{name}_phi.npy"), np.ascontiguousarray(oracle.interior(phi)))
        print(f"{name}: n_iter={nit} residual={res:.3e}")


if __name__ == "__main__":
    main()
    pp_main()
