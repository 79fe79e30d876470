"""Singular components labelled on the GPU (eafp_batch_components, and the
mixed solver's zero-mean projection that uses them) against scipy's
connected components of the same link graph, and the library's counters:
the FP64 drop-in and FP64 batches never label, a mixed solve labels each
upload once.

A component is singular when every row is a pure-Neumann row: aP equal to
the sum of its interior links within 1e-10 relative and no wall link.  The
test systems cut pockets out of a pure-Neumann system by removing the links
across a box's faces (each pocket and the rest are then singular), anchor
one pocket with a compliance term (not singular), and project the RHS to
zero mean on each singular component.
"""
import numpy as np
import pytest

import eafp
import oracle
from eafp import systems
from eafp._lib import COEF_BF16, EAFP_ERR_COMPONENTS, PRECOND_ZLINE

pytestmark = pytest.mark.gpu

NAMES = ("aW", "aE", "aS", "aN", "aB", "aT", "aP", "Su")


class Sys:
    def __init__(self, shape, arrs):
        self.shape = shape
        self.nr, self.nth, self.nz = shape
        self._a = arrs

    def padded(self, n):
        return self._a[n]

    def interior(self, n):
        return oracle.interior(self._a[n])


def expected_components(S):
    """(labels per interior cell (-1: not in a singular component), cell
    counts in order of first cell) from scipy's connected components."""
    import scipy.sparse as sp
    from scipy.sparse.csgraph import connected_components
    nr, nth, nz = S.shape
    a = {n: S.interior(n) for n in NAMES}
    idx = np.arange(nr * nth * nz).reshape(nr, nth, nz)
    rows, cols = [], []
    # E links (i, i+1), N links (j, j+1 periodic), T links (k, k+1): either side nonzero
    e = (a["aE"][:-1] != 0) | (a["aW"][1:] != 0)
    rows.append(idx[:-1][e]); cols.append(idx[1:][e])
    n_ = (a["aN"] != 0) | (np.roll(a["aS"], -1, axis=1) != 0)
    rows.append(idx[n_]); cols.append(np.roll(idx, -1, axis=1)[n_])
    t = (a["aT"][:, :, :-1] != 0) | (a["aB"][:, :, 1:] != 0)
    rows.append(idx[:, :, :-1][t]); cols.append(idx[:, :, 1:][t])
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    G = sp.coo_matrix((np.ones(r.size), (r, c)), shape=(idx.size, idx.size)).tocsr()
    ncomp, lab = connected_components(G, directed=False)
    linked = np.zeros(idx.size, dtype=bool)
    linked[r] = True
    linked[c] = True
    links = a["aS"] + a["aN"]
    links = links + np.pad(a["aW"][1:], ((1, 0), (0, 0), (0, 0))) + np.pad(a["aE"][:-1], ((0, 1), (0, 0), (0, 0)))
    links = links + np.pad(a["aB"][:, :, 1:], ((0, 0), (0, 0), (1, 0))) + \
        np.pad(a["aT"][:, :, :-1], ((0, 0), (0, 0), (0, 1)))
    wall = np.zeros(S.shape, dtype=bool)
    wall[0] |= a["aW"][0] != 0
    wall[-1] |= a["aE"][-1] != 0
    wall[:, :, 0] |= a["aB"][:, :, 0] != 0
    wall[:, :, -1] |= a["aT"][:, :, -1] != 0
    neumann = (~wall & (np.abs(a["aP"] - links) <= 1e-10 * np.abs(a["aP"]))).ravel()
    sing = np.ones(ncomp, dtype=bool)
    np.logical_and.at(sing, lab, linked & neumann)
    has_link = np.zeros(ncomp, dtype=bool)
    has_link[lab[linked]] = True
    sing &= has_link
    out = np.full(idx.size, -1)
    order = []
    for cell in range(idx.size):         # first cells in index order
        L = lab[cell]
        if linked[cell] and sing[L] and L not in order:
            order.append(L)
    counts = []
    for q, L in enumerate(order):
        m = (lab == L) & linked
        out[m] = q
        counts.append(int(m.sum()))
    return out.reshape(S.shape), counts


def pocket_system(boxes, anchored=(), seed=0):
    """Config-1 system with the links across each box's faces removed (aP
    lowered by what it loses, so every row stays pure Neumann), the boxes
    listed in `anchored` given a compliance term, and Su projected to zero
    mean on every singular component."""
    ps = systems.make_system(1, seed)
    arrs = {n: np.array(ps.padded(n), order="F") for n in NAMES}
    nr, nth, nz = ps.shape
    A = {n: arrs[n][2:-2, 2:-2, 2:-2] for n in NAMES}      # views
    for (i0, i1, j0, j1, k0, k1) in boxes:
        inside = np.zeros(ps.shape, dtype=bool)
        inside[i0:i1, j0:j1, k0:k1] = True
        for name, opp, axis in (("aE", "aW", 0), ("aN", "aS", 1), ("aT", "aB", 2)):
            nb_inside = np.roll(inside, -1, axis=axis)
            cut = inside != nb_inside
            if axis != 1:   # no link across the far wall
                sl = [slice(None)] * 3
                sl[axis] = -1
                cut[tuple(sl)] = False
            w = np.where(cut, A[name], 0.0)
            w_opp = np.roll(np.where(np.roll(cut, 1, axis=axis), A[opp], 0.0), 0, axis=axis)
            A["aP"] -= w
            A[name][cut] = 0.0
            A["aP"] -= w_opp
            A[opp][np.roll(cut, 1, axis=axis)] = 0.0
    for q in anchored:
        i0, i1, j0, j1, k0, k1 = boxes[q]
        A["aP"][i0, j0, k0] *= 1.001
    S = Sys(ps.shape, arrs)
    lab, counts = expected_components(S)
    su = A["Su"]
    for q in range(len(counts)):
        m = lab == q
        su[m] -= su[m].mean()
    return S, lab, counts


BOXES = [(1, 3, 2, 5, 1, 4), (6, 8, 10, 13, 5, 8), (9, 11, 18, 22, 9, 12), (3, 5, 20, 23, 11, 14),
         (7, 9, 4, 7, 2, 4)]


def test_components_match_scipy_and_project():
    """Three singular pockets, one anchored pocket and the rest of the mesh:
    the GPU's count and sizes equal scipy's, and the mixed result has zero
    mean on each singular component."""
    S, lab, counts = pocket_system(BOXES[:4], anchored=(1,))
    assert len(counts) == 4
    with eafp.Batch.from_systems([S]) as b:
        n, cells = b.components(0)
        print(f"case components pockets: gpu {n} {cells} | scipy {len(counts)} {counts}", flush=True)
        assert n == len(counts) and cells == counts
        st = b.solve_mixed(tol=1e-8, coef_format=COEF_BF16, precond=PRECOND_ZLINE)[0]
        phi = b.download(0)
    assert st["status"] == "converged"
    x = oracle.interior(phi)
    for q in range(len(counts)):
        m = lab == q
        assert abs(x[m].mean()) <= 1e-12 * np.abs(x[m]).max()
    phi_ref, _, _, _ = oracle.cg(S, max_iter=200000, tol=1e-12)
    err = oracle.gauge_error(x, oracle.interior(phi_ref), lab, len(counts))
    print(f"  inner={st['n_iter']} res={st['residual']:.2e} err={err:.2e}", flush=True)
    assert err < 1e-5


def test_more_than_four_singular_components_refused():
    """Five pockets and the rest: six singular components.  The labelling
    reports all six (and no sizes beyond four); the mixed solver refuses the
    system (EAFP_ERR_COMPONENTS), the FP64 solver still solves it, and a
    system without overflow in the same batch is labelled as well."""
    S, lab, counts = pocket_system(BOXES)
    ok = systems.make_system(1, 1)
    assert len(counts) == 6
    with eafp.Batch.from_systems([S, ok]) as b:
        n, cells = b.components(0)
        print(f"case components overflow: gpu {n} {cells} | scipy {len(counts)} {counts[:4]}", flush=True)
        assert n == 6 and cells == []
        n1, cells1 = b.components(1)
        lab1, counts1 = expected_components(ok)
        assert n1 == len(counts1) and cells1 == counts1[:4]
        with pytest.raises(eafp.EafpError) as e:
            b.solve_mixed(tol=1e-8, coef_format=COEF_BF16)
        assert e.value.rc == EAFP_ERR_COMPONENTS
        st = b.solve_fp64(tol=1e-8)
        assert st[0]["status"] == "converged" and st[1]["status"] == "converged"


@pytest.mark.parametrize("config,index,shape", [(1, 0, None), (3, 2, None), (5, 1, (24, 48, 32)),
                                                (2, 0, None)])
def test_generator_systems_components(config, index, shape):
    ps = systems.make_system(config, index, shape)
    lab, counts = expected_components(ps)
    with eafp.Batch.from_systems([ps]) as b:
        n, cells = b.components(0)
    print(f"case components config={config} shape={ps.shape}: gpu {n} {cells} | scipy {counts}", flush=True)
    assert n == len(counts) and cells == counts[:4]
    assert n == ps.ncomp


def test_dropin_and_fp64_batches_never_label():
    ps = systems.make_system(1)
    nr, nth, nz = ps.shape
    c0 = eafp.counters()
    phi = np.zeros((nr + 4, nth + 4, nz + 4), order="F")
    res, nit = eafp.cg_3d_mpi(*[ps.padded(n) for n in NAMES], phi, eafp.Mesh(nr, nth, nz), 5000, 1e-8)
    c1 = eafp.counters()
    print(f"case counters drop-in: {c0} -> {c1}", flush=True)
    assert res < 1e-8
    assert c1["uploads"] == c0["uploads"] + 1 and c1["labelled"] == c0["labelled"]
    with eafp.Batch.from_systems([ps, systems.make_system(1, 1)]) as b:
        b.solve_fp64(tol=1e-8)
        c2 = eafp.counters()
        assert c2["uploads"] == c1["uploads"] + 2 and c2["labelled"] == c1["labelled"]
        b.solve_mixed(tol=1e-8, coef_format=COEF_BF16)
        c3 = eafp.counters()
        assert c3["labelled"] == c2["labelled"] + 2
        b.solve_mixed(tol=1e-8, coef_format=COEF_BF16)
        c4 = eafp.counters()
        assert c4["labelled"] == c3["labelled"]
        assert c4["graph_captures"] == c3["graph_captures"]      # the same graph, replayed
        assert c4["graph_launches"] > c3["graph_launches"]
        b.upload(1, systems.make_system(1, 1))
        b.solve_mixed(tol=1e-8, coef_format=COEF_BF16)
        c5 = eafp.counters()
        assert c5["labelled"] == c4["labelled"] + 1
