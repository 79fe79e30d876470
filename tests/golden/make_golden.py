"""Writes the regression fixtures of tests/golden/ (run from the repository root:
``python tests/golden/make_golden.py``).

The reference holds no golden vectors or known-answer tests for its p' CG
(cg_3d_mpi, src/mod_solver_3d.f90:485-620), and its Fortran/MPI build cannot
be made in this image, so these fixtures are written by this repository's own
FP64 oracle (oracle/cg_oracle.c).  They pin two things against silent
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
