"""A p' system that exercises the wall-halo terms of cg_3d_mpi's first
matvec (src/mod_solver_3d.f90:509-513,593-604): nonzero wall links (aW at
i = 1, aE at i = nr, aB at k = 1, aT at k = nz, each added to its row's aP),
an initial phi with nonzero interior values, nonzero r/z halos and garbage
in its theta halo planes, and a nonzero RHS.  r0 = Su - A phi0 then reads
every r/z halo through a wall link."""
import numpy as np

from eafp import systems

NAMES = ("aW", "aE", "aS", "aN", "aB", "aT", "aP", "Su")


class FoldCase:
    def __init__(self, shape, arrs):
        self.shape = shape
        self.nr, self.nth, self.nz = shape
        self._a = arrs
        self.Su = arrs["Su"]

    def padded(self, n):
        return self._a[n]

    def interior(self, n):
        return self._a[n][2:-2, 2:-2, 2:-2]


def fold_case(shape=(20, 36, 28), seed=3):
    """(system, phi0): the workload-2 generator's system on `shape` with wall
    links added, and the initial phi."""
    ps = systems.make_system(2, 0, shape)
    rng = np.random.default_rng(seed)
    arrs = {n: np.array(ps.padded(n), order="F") for n in NAMES}
    I = (slice(2, -2),) * 3
    links = sum(np.abs(arrs[n][I]) for n in NAMES[:6])
    active = links > 0
    scale = float(np.median(links[active]))
    for name, sl in (("aW", (2, slice(2, -2), slice(2, -2))), ("aE", (-3, slice(2, -2), slice(2, -2))),
                     ("aB", (slice(2, -2), slice(2, -2), 2)), ("aT", (slice(2, -2), slice(2, -2), -3))):
        act = np.abs(arrs["aN"][sl]) + np.abs(arrs["aS"][sl]) > 0
        w = np.where(act, rng.uniform(0.2, 1.0, act.shape) * scale, 0.0)
        arrs[name][sl] = w
        arrs["aP"][sl] = arrs["aP"][sl] + w
    phi0 = np.asfortranarray(rng.normal(0.0, 1.0, arrs["aP"].shape))
    nth = shape[1]
    phi0[:, [0, 1, nth + 2, nth + 3], :] = 1e6        # theta halos: overwritten by the periodic copy
    return FoldCase(shape, arrs), phi0
