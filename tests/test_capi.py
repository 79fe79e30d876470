"""CPU checks of the C-ABI boundary: libeafp.so loads, exports every entry
point that include/eafp.h declares, and rejects invalid descriptors before
any GPU work (so these run without a GPU)."""
import ctypes
import os
import re

import numpy as np
import pytest

import eafp
from eafp import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    src = open(os.path.join(ROOT, "include", "eafp.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(eafp_[a-z0-9_]+)\s*\(", src)))


def test_header_symbols_exported():
    L = _lib.load()
    names = header_functions()
    assert set(names) == set(_lib.EXPORTS), names
    for n in names:
        assert hasattr(L, n), n
    assert L.eafp_version().startswith(b"eafp")


def test_struct_layouts_match_header():
    # eafp_desc: 7 ints; eafp_stats: double, 4 ints, 3 doubles; eafp_mixed_opts: double, 5 ints (+4 padding)
    assert ctypes.sizeof(_lib.Desc) == 28
    assert ctypes.sizeof(_lib.Stats) == 48
    assert ctypes.sizeof(_lib.MixedOpts) == 32
    assert _lib.MixedOpts.precond.offset == 24


def _arrays(nr=4, nth=6, nz=5):
    return [np.zeros((nr + 4, nth + 4, nz + 4), order="F") for _ in range(9)]


@pytest.mark.parametrize("field,value,rc", [
    ("is_parallel", 1, _lib.EAFP_ERR_UNSUPPORTED),
    ("halo", 1, _lib.EAFP_ERR_UNSUPPORTED),
    ("lbound", 0, _lib.EAFP_ERR_UNSUPPORTED),
    ("periodic_theta", 0, _lib.EAFP_ERR_UNSUPPORTED),
    ("nr", 0, _lib.EAFP_ERR_ARG),
])
def test_descriptor_rejected(field, value, rc):
    L = _lib.load()
    d = _lib.Desc(4, 6, 5, 2, -1, 0, 1)
    setattr(d, field, value)
    a = _arrays()
    res = ctypes.c_double(0.0)
    nit = ctypes.c_int(0)
    got = L.eafp_cg_solve(ctypes.byref(d), *[x.ctypes.data for x in a], 10, 1e-8, ctypes.byref(res), ctypes.byref(nit))
    assert got == rc


def test_null_pointer_rejected():
    L = _lib.load()
    d = _lib.Desc(4, 6, 5, 2, -1, 0, 1)
    a = _arrays()
    ptrs = [x.ctypes.data for x in a]
    ptrs[3] = None
    res = ctypes.c_double(0.0)
    nit = ctypes.c_int(0)
    assert L.eafp_cg_solve(ctypes.byref(d), *ptrs, 10, 1e-8, ctypes.byref(res), ctypes.byref(nit)) == _lib.EAFP_ERR_ARG


def test_python_mirror_maps_failure_to_nan():
    """Parallel meshes are refused; the mirror reports residual = NaN like the Fortran shim."""
    a = _arrays()
    m = eafp.Mesh(4, 6, 5, is_parallel=True)
    res, nit = eafp.cg_3d_mpi(*a[:8], a[8], m, 10, 1e-8)
    assert np.isnan(res) and nit == 0
    assert eafp.solver.last_status == _lib.EAFP_ERR_UNSUPPORTED


def test_layout_checked():
    a = _arrays()
    a[0] = np.zeros((8, 10, 9))  # C order
    m = eafp.Mesh(4, 6, 5)
    with pytest.raises(ValueError):
        eafp.cg_3d_mpi(*a[:8], a[8], m, 10, 1e-8)


def test_pp_struct_layouts():
    # eafp_pp_fields: 32 pointers; eafp_pp_config: 4 doubles, 4 ints
    assert ctypes.sizeof(_lib.PPFields) == 32 * 8
    assert ctypes.sizeof(_lib.PPConfig) == 48
    import oracle
    assert [f[0] for f in oracle.PPFields._fields_] == [f[0] for f in _lib.PPFields._fields_]
    assert ctypes.sizeof(oracle.PPConfig) == ctypes.sizeof(_lib.PPConfig)


def test_pp_assemble_rejects_bad_calls():
    """Argument checks run before any device call."""
    from eafp import fields
    F, cfg = fields.make_fields(4, 6, 5)
    with pytest.raises(_lib.EafpError) as e:
        eafp.pp_assemble(F, cfg, eafp.Mesh(4, 6, 5, is_parallel=True))
    assert e.value.rc == _lib.EAFP_ERR_UNSUPPORTED
    G = dict(F)
    del G["rho_l"]
    with pytest.raises(_lib.EafpError) as e:
        eafp.pp_assemble(G, cfg, eafp.Mesh(4, 6, 5))
    assert e.value.rc == _lib.EAFP_ERR_ARG
    with pytest.raises(_lib.EafpError) as e:
        eafp.pp_assemble(F, dict(cfg, dt=0.0), eafp.Mesh(4, 6, 5))
    assert e.value.rc == _lib.EAFP_ERR_ARG


@pytest.mark.parametrize("name,bad", [("rho_l", (8, 10, 8)), ("Ar", (8, 10)), ("outlet", (8, 10, 9)),
                                      ("r", (7,)), ("theta", (11,)), ("liq_room", (9, 10, 9))])
def test_pp_field_shapes_checked(name, bad):
    """Every field is checked against its padded shape before the library
    copies (nr+4)(ntheta+4)(nz+4) values (or n+4 coordinates, or the outlet
    layer) from it: a short array is a ValueError, never an over-read."""
    from eafp import fields
    shape = (4, 6, 5)
    F, cfg = fields.make_fields(*shape, masks=True)
    G = dict(F)
    G[name] = np.zeros(bad, dtype=F[name].dtype, order="F")
    m = eafp.Mesh(*shape)
    for call in (lambda: eafp.pp_assemble(G, cfg, m), lambda: eafp.pp_step(G, cfg, m, 0.3)):
        with pytest.raises(ValueError, match=name):
            call()
    pad = tuple(n + 4 for n in shape)
    ex = {n: np.zeros(pad, dtype=np.int32 if n.startswith("lk") else np.float64, order="F") for n in _lib.PP_EXPORTS}
    with pytest.raises(ValueError, match=name):
        eafp.pp_correct(G, cfg, m, 0.3, np.zeros(pad, order="F"), ex)
    with pytest.raises(ValueError, match="pp"):
        eafp.pp_correct(F, cfg, m, 0.3, np.zeros((8, 10, 8), order="F"), ex)
    ex["ac_z"] = np.zeros((8, 10, 8), order="F")
    with pytest.raises(ValueError, match="ac_z"):
        eafp.pp_correct(F, cfg, m, 0.3, np.zeros(pad, order="F"), ex)


def test_pp_assemble_noalloc_refuses_before_the_device():
    """The allocation-free assembly checks its arguments and its size limit
    before any device call: NULL arrays, a parallel mesh, and a mesh beyond
    EAFP_NOALLOC_MAX_CELLS padded cells (EAFP_ERR_SIZE)."""
    from eafp import fields
    shape = (4, 6, 5)
    F, cfg = fields.make_fields(*shape)
    f, c, keep = eafp.solver._pp_structs(F, cfg, eafp.Mesh(*shape))
    L = _lib.load()
    pad = np.zeros(tuple(n + 4 for n in shape), order="F")
    outs = [pad.ctypes.data] * 8
    big = eafp.solver.desc_of(eafp.Mesh(100, 100, 100))
    assert 104 ** 3 > _lib.NOALLOC_MAX_CELLS
    assert L.eafp_pp_assemble_noalloc(ctypes.byref(big), ctypes.byref(f), ctypes.byref(c), *outs,
                                      None, None, None, None) == _lib.EAFP_ERR_SIZE
    d = eafp.solver.desc_of(eafp.Mesh(*shape))
    assert L.eafp_pp_assemble_noalloc(ctypes.byref(d), ctypes.byref(f), ctypes.byref(c), *outs[:7], None,
                                      None, None, None, None) == _lib.EAFP_ERR_ARG
    par = eafp.solver.desc_of(eafp.Mesh(*shape, is_parallel=True))
    assert L.eafp_pp_assemble_noalloc(ctypes.byref(par), ctypes.byref(f), ctypes.byref(c), *outs,
                                      None, None, None, None) == _lib.EAFP_ERR_UNSUPPORTED
    del keep


def test_pp_post_solve_struct_layouts():
    # eafp_pp_exports: 9 pointers; eafp_pp_update: 13 pointers
    assert ctypes.sizeof(_lib.PPExports) == 9 * 8
    assert ctypes.sizeof(_lib.PPUpdate) == 13 * 8
    import oracle
    assert [f[0] for f in oracle.PPExports._fields_] == [f[0] for f in _lib.PPExports._fields_]


def test_pp_correct_rejects_bad_calls():
    """Argument checks of eafp_pp_correct run before any device call."""
    from eafp import fields
    shape = (4, 6, 5)
    F, cfg = fields.make_fields(*shape)
    pad = tuple(n + 4 for n in shape)
    ex = {n: np.zeros(pad, dtype=np.int32 if n.startswith("lk") else np.float64, order="F") for n in _lib.PP_EXPORTS}
    pp = np.zeros(pad, order="F")
    with pytest.raises(_lib.EafpError) as e:
        eafp.pp_correct(F, cfg, eafp.Mesh(*shape, is_parallel=True), 0.3, pp, ex)
    assert e.value.rc == _lib.EAFP_ERR_UNSUPPORTED
    G = dict(F)
    del G["aPz_g"]
    with pytest.raises(_lib.EafpError) as e:
        eafp.pp_correct(G, cfg, eafp.Mesh(*shape), 0.3, pp, ex)
    assert e.value.rc == _lib.EAFP_ERR_ARG
    # a NULL export array, through the C-ABI directly
    L = _lib.load()
    f, c, keep = eafp.solver._pp_structs(F, cfg, eafp.Mesh(*shape))
    vals = {n: a.ctypes.data for n, a in ex.items()}
    vals["ac_th"] = None
    exs = _lib.PPExports(**vals)
    io_arrays = {n: np.zeros(pad, dtype=np.int32 if n.startswith("lk") else np.float64, order="F")
                 for n in _lib.PP_UPDATE}
    io = _lib.PPUpdate(**{n: a.ctypes.data for n, a in io_arrays.items()})
    d = eafp.solver.desc_of(eafp.Mesh(*shape))
    rc = L.eafp_pp_correct(ctypes.byref(d), ctypes.byref(f), ctypes.byref(c), 0.3, pp.ctypes.data, ctypes.byref(exs),
                           ctypes.byref(io), 0)
    assert rc == _lib.EAFP_ERR_ARG
    rc = L.eafp_pp_correct(ctypes.byref(d), ctypes.byref(f), ctypes.byref(c), 0.3, None, ctypes.byref(exs),
                           ctypes.byref(io), 0)
    assert rc == _lib.EAFP_ERR_ARG
    del keep


def test_tdma_rejects_bad_calls():
    """Argument checks of the line TDMA run before any device call; n_sweep = 0 is a no-op."""
    shape = (4, 6, 5)
    a = _arrays()
    with pytest.raises(_lib.EafpError) as e:
        eafp.tdma_3d_mpi(*a[:8], a[8], eafp.Mesh(*shape, is_parallel=True), 2)
    assert e.value.rc == _lib.EAFP_ERR_UNSUPPORTED
    with pytest.raises(_lib.EafpError) as e:
        eafp.tdma_3d_mpi(*a[:8], a[8], eafp.Mesh(*shape), -1)
    assert e.value.rc == _lib.EAFP_ERR_ARG
    phi = a[8].copy(order="F")
    eafp.tdma_3d_mpi(*a[:8], phi, eafp.Mesh(*shape), 0)
    np.testing.assert_array_equal(phi, a[8])
    L = _lib.load()
    d = (_lib.Desc * 2)(eafp.solver.desc_of(eafp.Mesh(*shape)), eafp.solver.desc_of(eafp.Mesh(*shape)))
    vals = {n: x.ctypes.data for n, x in zip(("aW", "aE", "aS", "aN", "aB", "aT", "aP", "Su", "phi"), a)}
    good = _lib.TdmaSystem(**vals)
    vals["aN"] = None
    tab = (_lib.TdmaSystem * 2)(good, _lib.TdmaSystem(**vals))
    assert L.eafp_tdma_solve_batch(2, d, tab, 1, 0) == _lib.EAFP_ERR_ARG
    assert L.eafp_tdma_solve_batch(0, d, tab, 1, 0) == _lib.EAFP_ERR_ARG


def test_pp_step_struct_layout():
    # eafp_pp_step_opts: 2 ints, 2 doubles, then an eafp_mixed_opts (32 bytes) at offset 24
    assert ctypes.sizeof(_lib.PPStepOpts) == 56
    assert _lib.PPStepOpts.mixed.offset == 24
    assert _lib.PPStepOpts.alpha_p.offset == 16


def test_pp_step_rejects_bad_calls():
    """Argument checks of the fused p' step run before any device call."""
    from eafp import fields
    shape = (4, 6, 5)
    F, cfg = fields.make_fields(*shape)
    with pytest.raises(_lib.EafpError) as e:
        eafp.pp_step(F, cfg, eafp.Mesh(*shape, is_parallel=True), 0.3)
    assert e.value.rc == _lib.EAFP_ERR_UNSUPPORTED
    G = dict(F)
    del G["T_old_g"]
    with pytest.raises(_lib.EafpError) as e:
        eafp.pp_step(G, cfg, eafp.Mesh(*shape), 0.3)
    assert e.value.rc == _lib.EAFP_ERR_ARG
    with pytest.raises(_lib.EafpError) as e:
        eafp.pp_step(F, dict(cfg, dt=-1.0), eafp.Mesh(*shape), 0.3)
    assert e.value.rc == _lib.EAFP_ERR_ARG
    for bad in (dict(tol=0.0), dict(max_steps=0), dict(interval=-2), dict(coef_format=7)):
        with pytest.raises(_lib.EafpError) as e:
            eafp.pp_step(F, cfg, eafp.Mesh(*shape), 0.3, solver="mixed", mixed=bad)
        assert e.value.rc == _lib.EAFP_ERR_ARG, bad
    with pytest.raises(ValueError):
        eafp.pp_step(F, cfg, eafp.Mesh(*shape), 0.3, solver="sor")
    # an unknown solver code, a NULL pp and a NULL io array, through the C-ABI directly
    L = _lib.load()
    f, c, keep = eafp.solver._pp_structs(F, cfg, eafp.Mesh(*shape))
    pad = tuple(n + 4 for n in shape)
    io_arrays = {n: np.zeros(pad, dtype=np.int32 if n.startswith("lk") else np.float64, order="F")
                 for n in _lib.PP_UPDATE}
    io = _lib.PPUpdate(**{n: a.ctypes.data for n, a in io_arrays.items()})
    pp = np.zeros(pad, order="F")
    d = eafp.solver.desc_of(eafp.Mesh(*shape))
    mixed = _lib.MixedOpts(1e-8, 32, 100, _lib.COEF_BF16, 0)

    def call(o, pp_ptr, io_s):
        return L.eafp_pp_step(ctypes.byref(d), ctypes.byref(f), ctypes.byref(c), ctypes.byref(o), pp_ptr,
                              ctypes.byref(io_s), None, None, 0)

    assert call(_lib.PPStepOpts(2, 50, 1e-8, 0.3, mixed), pp.ctypes.data, io) == _lib.EAFP_ERR_ARG
    assert call(_lib.PPStepOpts(0, 50, 1e-8, 0.3, mixed), None, io) == _lib.EAFP_ERR_ARG
    vals = {n: a.ctypes.data for n, a in io_arrays.items()}
    vals["uth_g"] = None
    assert call(_lib.PPStepOpts(0, 50, 1e-8, 0.3, mixed), pp.ctypes.data, _lib.PPUpdate(**vals)) == _lib.EAFP_ERR_ARG
    # valid arguments reach the device check: no device here, or a GPU run
    rc = call(_lib.PPStepOpts(0, 50, 1e-8, 0.3, mixed), pp.ctypes.data, io)
    assert rc in (_lib.EAFP_OK, _lib.EAFP_ERR_NO_DEVICE)
    del keep
