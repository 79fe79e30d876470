"""p' CG solves/s at 60x120x84 to 1e-8 relative residual: mixed precision
(batched) vs the plain FP64 GPU CG, with the roofline of the dominant kernel
and the CPU oracle beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--systems B]

One step = one batched mixed-precision solve (FP32 inner CG on BF16
symmetric face weights + FP64 refinement, setup pass included) of B seeded
60x120x84 campaign-variant systems already resident in HBM, from phi = 0 to
a true relative residual of 1e-8 on every system.  `value` = B*K / elapsed
seconds.

Beside it, on the same GPU:
- the same batch with the Jacobi inner preconditioner (the reference's);
- the same batch solved by the batched plain FP64 Jacobi-PCG (the cg_3d_mpi
  algorithm): the like-for-like denominator of the speedup;
- uploads from device-resident arrays, the singular-component labelling and
  the solve after them (upload_plus_solve_solves_per_s);
- parity of every result: the relative p' error of each of the B systems
  against an FP64 Jacobi-PCG run to 1e-12, and its true and L1 continuity
  residuals recomputed on the host;
- the plain FP64 GPU CG and the mixed solver on the single 60x120x84
  production system, the drop-in on host arrays;
- the CPU restatement of cg_3d_mpi (OpenMP and 1 core) on the production
  system;
- the p' assembly, the fused p' step, the line TDMA, and the other workloads
  (256 ragged systems; one 120x240x168 system), each with its parity.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "hornofusion-full3d_amd"))

import numpy as np  # noqa: E402

METRIC = "p' CG solves/s @60×120×84 to 1e-8 rel res (mixed vs FP64); HBM GB/s"
HBM_PEAK_GBS = 8000.0
# Algorithmic HBM bytes per cell of one launch (see the kernel headers):
BYTES_MSPMV = {"fp32": 32, "bf16": 24, "fp16": 28}   # z, p_old, 3 weights and e (+FP32 scale); p, q
BYTES_MUPDATE = 32                                   # d, r, p, q, 1/D (z-line: l, 1/d); d, r, z
BYTES_FP64_ITER = 152                                # k64_spmv 88 + k64_update 64
CALIB_BYTES = 1 << 30                                # calibration copy for the HBM counters


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def device_sync():
    """hipDeviceSynchronize (the library's calls already return synchronised)."""
    try:
        hip = ctypes.CDLL("libamdhip64.so")
        hip.hipDeviceSynchronize()
    except OSError:
        pass


def calibration_copy(nbytes):
    """A plain device-to-device copy of known size: the reference point that
    scales the HBM counters (FETCH_SIZE/WRITE_SIZE) of the solver kernels."""
    hip = ctypes.CDLL("libamdhip64.so")
    a = ctypes.c_void_p()
    b = ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) or hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, ctypes.c_size_t(nbytes))
    hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3)   # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)


def spmv_kernel_name(info, coef):
    f = {"fp32": 0, "bf16": 1, "fp16": 2}[coef]
    cpt = info["tile_cells"] // 256
    if info["column_splits"] > 0:
        return f"k_mspmv_col<{f}>"
    return f"k_mspmv4<{f}, {cpt // 4}>" if cpt >= 4 else f"k_mspmv<{f}, {cpt}>"


def update_kernel_name(info, zline=False):
    if zline:
        return "k_mupdate_line"
    cpt = info["tile_cells"] // 256
    return f"k_mupdate4<{cpt // 4}>" if cpt >= 4 else f"k_mupdate<{cpt}>"


def read_traffic(kernel, workload):
    """HBM bytes per launch measured with rocprofv3 counters (profiles/), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)
        return rec.get(workload, {}).get(kernel, {}).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


def _counter_rows(path):
    """(kernel name, counter name, value) of every row of the rocprofv3
    counter-collection CSV files under `path` (a file or a directory)."""
    import csv
    import glob
    files = [path] if os.path.isfile(path) else sorted(
        glob.glob(os.path.join(path, "**", "*counter_collection.csv"), recursive=True))
    if not files:
        raise FileNotFoundError(f"no *counter_collection.csv under {path}")
    rows = []
    for fn in files:
        with open(fn, newline="") as f:
            for r in csv.DictReader(f):
                rows.append((r["Kernel_Name"], r["Counter_Name"], float(r["Counter_Value"])))
    return rows


def kernel_key(kname):
    """The solver kernel a profiled dispatch belongs to (None: other)."""
    for key in ("k64_spmv", "k64_update", "k_mspmv", "k_mupdate"):
        if key in kname:
            return key
    return "copy" if "copyBuffer" in kname else None


def derive_traffic(fetch_path, write_path, workload, algo):
    """HBM bytes per launch of the solver kernels from two rocprofv3 --pmc
    runs of `bench.py --kernels-only` (FETCH_SIZE, then WRITE_SIZE, in KiB
    per dispatch).  Each counter is scaled by what it reads for the
    calibration copy of CALIB_BYTES (the runtime's device-to-device copy
    kernel), so the result is in bytes the HBM actually moved.  `algo` maps
    each kernel key to its algorithmic bytes per launch."""
    per = {}
    for path, counter in ((fetch_path, "FETCH_SIZE"), (write_path, "WRITE_SIZE")):
        for kname, cname, val in _counter_rows(path):
            if cname != counter:
                continue
            key = kernel_key(kname)
            if key:
                per.setdefault((key, counter), []).append(val * 1024.0)
    calib = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        v = per.get(("copy", counter))
        if not v:
            raise RuntimeError(f"calibration copy not found in the {counter} run")
        calib[counter] = CALIB_BYTES / max(v)     # the 1 GiB copy is the largest copy dispatch
    rec = {"calibration": {"copy_bytes": CALIB_BYTES, "fetch_factor": round(calib["FETCH_SIZE"], 4),
                           "write_factor": round(calib["WRITE_SIZE"], 4)}}
    for k in algo:
        f = per.get((k, "FETCH_SIZE"), [])
        w = per.get((k, "WRITE_SIZE"), [])
        if not f or not w:
            continue
        fb = float(np.mean(f)) * calib["FETCH_SIZE"]
        wb = float(np.mean(w)) * calib["WRITE_SIZE"]
        rec[k] = {"bytes_per_launch": round(fb + wb), "fetch_bytes": round(fb), "write_bytes": round(wb),
                  "algorithmic_bytes": algo[k], "traffic_over_algorithmic": round((fb + wb) / algo[k], 4),
                  "dispatches": [len(f), len(w)]}
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as fh:
            allrec = json.load(fh)
    except (OSError, ValueError):
        allrec = {}
    allrec[workload] = rec
    with open(path, "w") as fh:
        json.dump(allrec, fh, indent=1, sort_keys=True)
        fh.write("\n")
    return rec


def measure_workload(eafp, syss, fmt, coef, interval, label, precond=0):
    """Mixed (batched) and plain FP64 GPU solves of one workload to 1e-8:
    solves/s, iterations, the iteration kernels' algorithmic GB/s, and the
    parity of every mixed result against FP64 references (tol 1e-12).  The
    z-line preconditioner is used when asked for and every nz allows it."""
    ncells = sum(p.ncells for p in syss)
    b = eafp.Batch.from_systems(syss)
    if precond and not all(p.shape[2] <= 1024 for p in syss):
        precond = 0
    rec = {"workload": label, "systems": len(syss), "cells": ncells, "coef_format": coef,
           "precond": "zline" if precond else "jacobi"}
    msolver = eafp.Batch.SOLVER_MIXED_ZLINE if precond else eafp.Batch.SOLVER_MIXED
    b.autotune(msolver, fmt, reps=3)
    b.solve_mixed(tol=1e-8, interval=interval, coef_format=fmt, precond=precond)
    t = time.perf_counter()
    st = b.solve_mixed(tol=1e-8, interval=interval, coef_format=fmt, precond=precond)
    tm = time.perf_counter() - t
    xs = [b.download(s) for s in range(len(syss))]
    ms_sp, ms_up = b.time_kernels(msolver, fmt, reps=10)
    rec["mixed"] = {"solves_per_s": round(len(syss) / tm, 3), "seconds": round(tm, 4),
                    "all_converged": all(x["status"] == "converged" for x in st),
                    "inner_iters_max": max(x["n_iter"] for x in st),
                    "max_true_rel_residual": max(x["residual"] for x in st),
                    "launch_shape": b.info(),
                    "k_mspmv_GBps": round(BYTES_MSPMV[coef] * ncells / (ms_sp * 1e-3) / 1e9, 1),
                    "k_mupdate_GBps": round(BYTES_MUPDATE * ncells / (ms_up * 1e-3) / 1e9, 1)}
    b.autotune(eafp.Batch.SOLVER_FP64, reps=3)
    t = time.perf_counter()
    st64 = b.solve_fp64(200000, 1e-8)
    t64 = time.perf_counter() - t
    ms64 = b.time_kernels(eafp.Batch.SOLVER_FP64, reps=10)
    rec["fp64"] = {"solves_per_s": round(len(syss) / t64, 3), "seconds": round(t64, 4),
                   "all_converged": all(x["status"] == "converged" for x in st64),
                   "n_iter_max": max(x["n_iter"] for x in st64),
                   "k64_spmv_GBps": round(88 * ncells / (ms64[0] * 1e-3) / 1e9, 1),
                   "k64_update_GBps": round(64 * ncells / (ms64[1] * 1e-3) / 1e9, 1)}
    rec["speedup_mixed_vs_fp64"] = round(t64 / tm, 3)
    refs, stref = fp64_references(eafp, syss, batch=b)
    rec["parity"] = parity_summary(syss, xs, refs)
    rec["parity"]["fp64_reference_converged"] = all(x["status"] == "converged" for x in stref)
    b.close()
    return rec


def measure_assembly(eafp, shape=(60, 120, 84), reps=20):
    """The p' system assembled on the GPU (eafp_pp_assemble, with the
    liquid's face exports) and the post-solve consumers (eafp_pp_correct)
    on two-phase fields at 60x120x84: device-resident (every array already
    in HBM, the case of a GPU-resident caller) and, for the assembly, from
    host arrays (PCIe copies included), beside the oracle's single-core
    times."""
    import ctypes
    from eafp import _lib, fields
    F, cfg = fields.make_fields(*shape, seed=11, masks=True)
    m = eafp.Mesh(*shape)
    eafp.pp_assemble(F, cfg, m)
    t = time.perf_counter()
    for _ in range(reps):
        out, res, flux = eafp.pp_assemble(F, cfg, m)
    t_host = (time.perf_counter() - t) / reps
    # device-resident: every input and output in HBM
    hip = ctypes.CDLL("libamdhip64.so")
    bufs = []

    def dev(a=None, nbytes=None):
        ptr = ctypes.c_void_p()
        nb = a.nbytes if a is not None else nbytes
        _lib.check(hip.hipMalloc(ctypes.byref(ptr), ctypes.c_size_t(nb)), "hipMalloc")
        if a is not None:
            _lib.check(hip.hipMemcpy(ptr, ctypes.c_void_p(a.ctypes.data), ctypes.c_size_t(nb), 1), "hipMemcpy")
        bufs.append(ptr)
        return ptr.value

    vals = {}
    for n in _lib.PP_FIELDS:
        a = F.get(n)
        vals[n] = None if a is None else dev(np.asfortranarray(a, dtype=np.int32 if n in _lib.PP_INT else np.float64))
    npad = (shape[0] + 4) * (shape[1] + 4) * (shape[2] + 4)
    outs = [dev(nbytes=8 * npad) for _ in range(8)]
    exs = _lib.PPExports(**{n: dev(nbytes=(4 if n.startswith("lk") else 8) * npad) for n in _lib.PP_EXPORTS})
    f = _lib.PPFields(**vals)
    c = _lib.PPConfig(cfg["dt"], cfg["rho_gas"], cfg["T_ambient"], cfg["beta_expansion"],
                      cfg["gas_in_poisson"], cfg["solve_multiphase"], cfg["gas_compressibility"], 0)
    L = _lib.load()
    d = eafp.solver.desc_of(m)
    r_ = ctypes.c_double()
    fl = ctypes.c_double()
    call = lambda: _lib.check(L.eafp_pp_assemble(ctypes.byref(d), ctypes.byref(f), ctypes.byref(c), *outs,
                                                  ctypes.byref(r_), ctypes.byref(fl), ctypes.byref(exs), 1),
                              "eafp_pp_assemble")
    call()
    t = time.perf_counter()
    for _ in range(reps):
        call()
    t_dev = (time.perf_counter() - t) / reps
    # post-solve consumers on a smooth correction field (theta-periodic)
    nr, nth, nz = shape
    ii = np.arange(-1, nr + 3)[:, None, None]
    jj = np.arange(-1, nth + 3)[None, :, None]
    kk = np.arange(-1, nz + 3)[None, None, :]
    pp = np.asfortranarray(50.0 * np.sin(0.1 * ii) * np.cos(2 * np.pi * jj / nth) + 3.0 * kk)
    upd = {n: dev(nbytes=(4 if n.startswith("lk") else 8) * npad) for n in _lib.PP_UPDATE[:6]}
    upd.update({n: dev(np.asfortranarray(F[n], dtype=np.float64)) for n in _lib.PP_UPDATE[6:]})
    io = _lib.PPUpdate(**upd)
    pp_d = dev(pp)
    corr = lambda: _lib.check(L.eafp_pp_correct(ctypes.byref(d), ctypes.byref(f), ctypes.byref(c), 0.3, pp_d,
                                                 ctypes.byref(exs), ctypes.byref(io), 1), "eafp_pp_correct")
    corr()
    t = time.perf_counter()
    for _ in range(reps):
        corr()
    t_corr = (time.perf_counter() - t) / reps
    for ptr in bufs:
        hip.hipFree(ptr)
    rec = {"workload": "two-phase 60x120x84 fields (bath, gas, scrap band, workspace masks) -> aW..aP, Su, "
                       "face exports; then fluxes, velocities, p",
           "device_resident_us_per_call": round(t_dev * 1e6, 1), "host_arrays_ms_per_call": round(t_host * 1e3, 3),
           "post_solve_device_resident_us_per_call": round(t_corr * 1e6, 1),
           "continuity_residual": res}
    sys.path.insert(0, ROOT)
    import oracle
    oracle.lib()
    t = time.perf_counter()
    _, _, _, ex = oracle.pp_assemble(F, cfg, shape, exports=True)
    rec["cpu_oracle_ms_1_core"] = round((time.perf_counter() - t) * 1e3, 2)
    t = time.perf_counter()
    oracle.pp_correct(F, cfg, shape, 0.3, pp, ex)
    rec["post_solve_cpu_oracle_ms_1_core"] = round((time.perf_counter() - t) * 1e3, 2)
    return rec


def measure_step(eafp, shape=(60, 120, 84), reps=5):
    """One whole p' step (eafp_pp_step: assembly, solve from pp = 0,
    post-solve steps) on two-phase fields at 60x120x84, device-resident:
    with the reference's defaults (cfg%max_inner_pres = 50 FP64 CG
    iterations, tol 1e-8), and solved to 1e-8 by the FP64 CG and by the
    mixed solver (BF16; Jacobi and z-line inner preconditioners).  Beside it the same 50-iteration step from host
    arrays, fused and as the three unfused calls, and the oracle's three
    steps on one host core."""
    import ctypes
    from eafp import _lib, fields
    F, cfg = fields.make_fields(*shape, seed=11, masks=True)
    m = eafp.Mesh(*shape)
    hip = ctypes.CDLL("libamdhip64.so")
    bufs = []

    def dev(a=None, nbytes=None):
        ptr = ctypes.c_void_p()
        nb = a.nbytes if a is not None else nbytes
        _lib.check(hip.hipMalloc(ctypes.byref(ptr), ctypes.c_size_t(nb)), "hipMalloc")
        if a is not None:
            _lib.check(hip.hipMemcpy(ptr, ctypes.c_void_p(a.ctypes.data), ctypes.c_size_t(nb), 1), "hipMemcpy")
        bufs.append(ptr)
        return ptr.value

    vals = {}
    for n in _lib.PP_FIELDS:
        a = F.get(n)
        vals[n] = None if a is None else dev(np.asfortranarray(a, dtype=np.int32 if n in _lib.PP_INT else np.float64))
    npad = (shape[0] + 4) * (shape[1] + 4) * (shape[2] + 4)
    f = _lib.PPFields(**vals)
    c = _lib.PPConfig(cfg["dt"], cfg["rho_gas"], cfg["T_ambient"], cfg["beta_expansion"],
                      cfg["gas_in_poisson"], cfg["solve_multiphase"], cfg["gas_compressibility"], 0)
    # io arrays apart from the fields: every call assembles from the same inputs
    upd = {n: dev(nbytes=(4 if n.startswith("lk") else 8) * npad) for n in _lib.PP_UPDATE[:6]}
    upd.update({n: dev(np.asfortranarray(F[n], dtype=np.float64)) for n in _lib.PP_UPDATE[6:]})
    io = _lib.PPUpdate(**upd)
    pp_d = dev(nbytes=8 * npad)
    L = _lib.load()
    d = eafp.solver.desc_of(m)
    mixed = _lib.MixedOpts(1e-8, 32, 2000, _lib.COEF_BF16, 0, _lib.PRECOND_JACOBI)
    mixed_zl = _lib.MixedOpts(1e-8, 32, 2000, _lib.COEF_BF16, 0, _lib.PRECOND_ZLINE)
    rec = {"workload": "one p' step on two-phase 60x120x84 fields: assembly, solve from pp = 0, fluxes, velocities, p",
           "device_resident": {}}
    for label, o in (("fp64_50_iterations", _lib.PPStepOpts(_lib.STEP_FP64, 50, 1e-8, 0.3, mixed)),
                     ("fp64_to_1e-8", _lib.PPStepOpts(_lib.STEP_FP64, 100000, 1e-8, 0.3, mixed)),
                     ("mixed_bf16_to_1e-8", _lib.PPStepOpts(_lib.STEP_MIXED, 0, 0.0, 0.3, mixed)),
                     ("mixed_bf16_zline_to_1e-8", _lib.PPStepOpts(_lib.STEP_MIXED, 0, 0.0, 0.3, mixed_zl))):
        res = ctypes.c_double()
        st = _lib.Stats()
        call = lambda: _lib.check(L.eafp_pp_step(ctypes.byref(d), ctypes.byref(f), ctypes.byref(c), ctypes.byref(o),
                                                 pp_d, ctypes.byref(io), ctypes.byref(res), ctypes.byref(st), 1),
                                  "eafp_pp_step")
        call()
        t = time.perf_counter()
        for _ in range(reps):
            call()
        dt = (time.perf_counter() - t) / reps
        rec["device_resident"][label] = {"ms_per_step": round(dt * 1e3, 3), "steps_per_s": round(1.0 / dt, 2),
                                         "n_iter": st.n_iter, "n_refine": st.n_refine,
                                         "solver_residual": st.residual, "true_rel_l2": st.true_rel_l2}
        log(f"pp_step {label}: {dt * 1e3:.2f} ms, n_iter {st.n_iter}, true rel {st.true_rel_l2:.2e}")
    for ptr in bufs:
        hip.hipFree(ptr)
    # host arrays, 50 iterations: fused, then the three unfused calls
    eafp.pp_step(F, cfg, m, 0.3, max_iter=50)
    t = time.perf_counter()
    for _ in range(reps):
        eafp.pp_step(F, cfg, m, 0.3, max_iter=50)
    rec["host_arrays_fused_ms_per_step"] = round((time.perf_counter() - t) / reps * 1e3, 3)
    names = ("aW", "aE", "aS", "aN", "aB", "aT", "aP", "Su")

    def unfused():
        out, _, _, ex = eafp.pp_assemble(F, cfg, m, exports=True)
        pp = np.zeros_like(out["aP"], order="F")
        eafp.cg_3d_mpi(*[out[n] for n in names], pp, m, 50, 1e-8)
        eafp.pp_correct(F, cfg, m, 0.3, pp, ex)
    unfused()
    t = time.perf_counter()
    for _ in range(reps):
        unfused()
    rec["host_arrays_three_calls_ms_per_step"] = round((time.perf_counter() - t) / reps * 1e3, 3)
    sys.path.insert(0, ROOT)
    import oracle
    oracle.lib()
    t = time.perf_counter()
    ref, _, _, ex = oracle.pp_assemble(F, cfg, shape, exports=True)

    class S:
        pass
    ps = S()
    ps.shape = shape
    ps.padded = lambda n: ref[n]
    pp_ref, _, _, _ = oracle.cg(ps, max_iter=50, tol=1e-8)
    oracle.pp_correct(F, cfg, shape, 0.3, pp_ref, ex)
    cpu = time.perf_counter() - t
    rec["cpu_oracle_ms_per_step_50_iterations_1_core"] = round(cpu * 1e3, 1)
    rec["speedup_device_resident_vs_cpu"] = round(cpu * 1e3 / rec["device_resident"]["fp64_50_iterations"]["ms_per_step"], 1)
    return rec


def measure_tdma(eafp, batches=(1, 8, 64), n_sweep=3, reps=5):
    """The line TDMA (eafp_tdma_solve_batch, device-resident) on the 60x120x84
    production p' system: one system, the 8 equations of a two-phase time
    step side by side, and 64 systems; beside the oracle's sequential
    tdma_3d_mpi on one host core (the reference's smoother is serial per
    rank).  Sweep time = call time / n_sweep; the pivot pass (once per call)
    is included."""
    import ctypes
    from eafp import _lib, systems
    ps = systems.make_system(2)
    shape = ps.shape
    names = ("aW", "aE", "aS", "aN", "aB", "aT", "aP", "Su")
    arrs = [np.asfortranarray(ps.padded(n)) for n in names]
    npad = arrs[0].size
    hip = ctypes.CDLL("libamdhip64.so")
    L = _lib.load()
    rec = {"workload": f"tdma_3d_mpi, {n_sweep} sweeps of the 60x120x84 production p' system (604,800 cells), "
                       "device-resident; B independent copies solved side by side",
           "n_sweep": n_sweep, "batches": []}
    for B in batches:
        bufs = []

        def dev(a):
            ptr = ctypes.c_void_p()
            _lib.check(hip.hipMalloc(ctypes.byref(ptr), ctypes.c_size_t(a.nbytes)), "hipMalloc")
            _lib.check(hip.hipMemcpy(ptr, ctypes.c_void_p(a.ctypes.data), ctypes.c_size_t(a.nbytes), 1), "hipMemcpy")
            bufs.append(ptr)
            return ptr.value

        zero = np.zeros(npad)
        tab = (_lib.TdmaSystem * B)()
        for q in range(B):
            tab[q] = _lib.TdmaSystem(*[dev(a) for a in arrs], dev(zero))
        descs = (_lib.Desc * B)(*[eafp.solver.desc_of(eafp.Mesh(*shape)) for _ in range(B)])
        call = lambda: _lib.check(L.eafp_tdma_solve_batch(B, descs, tab, n_sweep, 1), "eafp_tdma_solve_batch")
        call()
        t = time.perf_counter()
        for _ in range(reps):
            call()
        dt = (time.perf_counter() - t) / reps
        for ptr in bufs:
            hip.hipFree(ptr)
        sweep_ms = dt / n_sweep * 1e3
        # unique bytes per cell and pass: the six arrays a pass reads (Su, four off-line
        # coefficients, the line's minus coefficient), pivot and c', phi read and written,
        # the forward scratch written and read: 96 B; three passes per sweep
        gbps = B * ps.ncells * 3 * 96 / (sweep_ms * 1e-3) / 1e9
        rec["batches"].append({"systems": B, "ms_per_call": round(dt * 1e3, 3), "ms_per_sweep": round(sweep_ms, 3),
                               "system_sweeps_per_s": round(B / (sweep_ms * 1e-3), 1),
                               "algorithmic_GBps": round(gbps, 1)})
    sys.path.insert(0, ROOT)
    import oracle
    oracle.lib()
    coef = dict(zip(names, arrs))
    t = time.perf_counter()
    oracle.tdma_3d(coef, np.zeros_like(arrs[0], order="F"), shape, 1)
    cpu = time.perf_counter() - t
    rec["cpu_oracle_ms_per_sweep_1_core"] = round(cpu * 1e3, 2)
    rec["speedup_single_vs_cpu"] = round(cpu * 1e3 / rec["batches"][0]["ms_per_sweep"], 1)
    return rec


def measure_upload(eafp, batch, syss, fmt, interval, precond):
    """Every system of `batch` uploaded again from device-resident padded
    arrays (eafp_batch_upload with arrays_on_device = 1: packed where they
    lie), then the first mixed solve after it, which labels the singular
    components of each system on the GPU (timed apart through
    eafp_batch_components) before solving."""
    from eafp import _lib
    hip = ctypes.CDLL("libamdhip64.so")
    bufs = []
    ptrs = []
    for ps in syss:
        row = []
        for n in NAMES8:
            a = np.asfortranarray(ps.padded(n))
            p = ctypes.c_void_p()
            _lib.check(hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(a.nbytes)), "hipMalloc")
            _lib.check(hip.hipMemcpy(p, ctypes.c_void_p(a.ctypes.data), ctypes.c_size_t(a.nbytes), 1), "hipMemcpy")
            bufs.append(p)
            row.append(p.value)
        ptrs.append(row)
    device_sync()
    t = time.perf_counter()
    for s, row in enumerate(ptrs):
        batch.upload_device(s, row)
    t_up = time.perf_counter() - t
    t = time.perf_counter()
    ns, _ = batch.components(0)
    t_lab = time.perf_counter() - t
    t = time.perf_counter()
    st = batch.solve_mixed(tol=1e-8, interval=interval, coef_format=fmt, precond=precond)
    t_solve = time.perf_counter() - t
    for p in bufs:
        hip.hipFree(p)
    B = len(syss)
    return {"arrays_on_device": 1, "upload_s": round(t_up, 4), "label_s": round(t_lab, 4),
            "solve_s": round(t_solve, 4), "all_converged": all(x["status"] == "converged" for x in st),
            "singular_components_system0": ns,
            "upload_plus_solve_solves_per_s": round(B / (t_up + t_lab + t_solve), 3)}


def rel_err(x, xr, ps):
    from eafp.systems import project_zero_mean
    px = project_zero_mean(x, ps.labels, ps.ncomp)
    pr = project_zero_mean(xr, ps.labels, ps.ncomp)
    return float(np.linalg.norm(px - pr) / max(np.linalg.norm(pr), 1e-300))


NAMES8 = ("aW", "aE", "aS", "aN", "aB", "aT", "aP", "Su")


def host_residuals(ps, phi):
    """FP64 residual of a downloaded phi, recomputed on the host with numpy
    (independent of the library): (||b - A x||_2 / ||b||_2, sum|b - A x| /
    sum|b|, the L1 form of res_cont in src/mod_pressure_3d.f90:293-306), with
    A x as the reference's matvec forms it (phi's theta halos hold their
    periodic copies after a download, its r/z halos what was uploaded)."""
    a = {n: ps.padded(n) for n in NAMES8}
    c = (slice(2, -2),) * 3
    x = phi
    ax = (a["aP"][c] * x[c] - a["aW"][c] * x[1:-3, 2:-2, 2:-2] - a["aE"][c] * x[3:-1, 2:-2, 2:-2]
          - a["aS"][c] * x[2:-2, 1:-3, 2:-2] - a["aN"][c] * x[2:-2, 3:-1, 2:-2]
          - a["aB"][c] * x[2:-2, 2:-2, 1:-3] - a["aT"][c] * x[2:-2, 2:-2, 3:-1])
    b = a["Su"][c]
    r = b - ax
    return (float(np.linalg.norm(r) / max(np.linalg.norm(b), 1e-300)),
            float(np.abs(r).sum() / max(np.abs(b).sum(), 1e-300)))


PARITY_BAR = ("each result against an FP64 Jacobi-PCG run (the cg_3d_mpi algorithm, GPU, tol 1e-12) of the same "
              "system: relative p' error ||P(x - x_ref)||/||P x_ref|| with P the zero-mean projection of every "
              "singular component; true relative residual ||b - A x||/||b|| and L1 continuity residual "
              "sum|b - A x|/sum|b|, both recomputed on the host in FP64")


def fp64_references(eafp, syss, tol=1e-12, batch=None):
    """The FP64 reference solutions: every system solved by the GPU FP64
    Jacobi-PCG (the drop-in's algorithm) to `tol`.  Returns (padded phis,
    stats)."""
    own = batch is None
    b = eafp.Batch.from_systems(syss) if own else batch
    st = b.solve_fp64(400000, tol)
    xs = [b.download(s) for s in range(len(syss))]
    if own:
        b.close()
    return xs, st


def parity_summary(syss, xs, refs):
    """Per-system p' error against the FP64 references and host-recomputed
    residuals, summarised over the systems."""
    errs, l2s, l1s = [], [], []
    for ps, x, xr in zip(syss, xs, refs):
        errs.append(rel_err(x[2:-2, 2:-2, 2:-2], xr[2:-2, 2:-2, 2:-2], ps))
        l2, l1 = host_residuals(ps, x)
        l2s.append(l2)
        l1s.append(l1)
    return {"systems_checked": len(errs), "max_rel_err_p": max(errs), "mean_rel_err_p": float(np.mean(errs)),
            "max_true_rel_residual_host": max(l2s), "max_l1_continuity_residual_host": max(l1s),
            "mean_l1_continuity_residual_host": float(np.mean(l1s)),
            "all_below_1e-8": bool(max(l2s) < 1e-8)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--systems", type=int, default=64)
    ap.add_argument("--coef", choices=("fp32", "bf16", "fp16"), default="bf16")
    ap.add_argument("--interval", type=int, default=32)
    ap.add_argument("--precond", choices=("jacobi", "zline"), default="zline",
                    help="inner preconditioner of the mixed solver")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--no-jacobi", action="store_true",
                    help="skip the Jacobi-inner leg of the z-line bench (one solver graph per run, for tracing)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-reference", action="store_true", help="skip the FP64 legs and error checks")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the other workloads (256 ragged systems; one 120x240x168 system in BF16)")
    ap.add_argument("--tiling", default=None,
                    help="pin the launch shape 'cells_per_thread,blocks[,column_splits]' (no autotune)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="only the roofline leg: a 1 GiB device copy (counter calibration) and the timed "
                         "k_mspmv/k_mupdate launches, as measured for the bench line")
    ap.add_argument("--single", action="store_true",
                    help="with --kernels-only / --derive-traffic: the single 60x120x84 production system and the "
                         "FP64 drop-in's kernels (k64_spmv, k64_update) instead of the batch")
    ap.add_argument("--assembly-only", action="store_true",
                    help="only the p' assembly and post-solve leg (60x120x84 two-phase fields), for profiling")
    ap.add_argument("--step-only", action="store_true",
                    help="only the fused p' step leg (assembly + solve + post-solve, 60x120x84), for profiling")
    ap.add_argument("--tdma-only", action="store_true",
                    help="only the line-TDMA leg (60x120x84, 1/8/64 systems), for profiling")
    ap.add_argument("--derive-traffic", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="write profiles/hbm_traffic.json from the rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE "
                         "outputs of two `bench.py --kernels-only` runs with the same options; no GPU use")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (--gpus {args.gpus} requested); "
              "the batched p' solve is sharded by system across GPUs in a multi-GPU build, which this is not.",
              file=sys.stderr)
        return 2

    B = args.systems
    zline = args.precond == "zline"
    precond = 1 if zline else 0
    workload = f"batch{B}_60x120x84_{args.coef}" + ("_zline" if zline else "")
    if args.derive_traffic:
        if args.single:
            n1 = 60 * 120 * 84
            workload, algo = "single_60x120x84_fp64", {"k64_spmv": 88 * n1, "k64_update": 64 * n1}
        else:
            nb = B * 60 * 120 * 84
            algo = {"k_mspmv": BYTES_MSPMV[args.coef] * nb, "k_mupdate": BYTES_MUPDATE * nb}
        rec = derive_traffic(args.derive_traffic[0], args.derive_traffic[1], workload, algo)
        print(json.dumps({workload: rec}, indent=1), flush=True)
        return 0

    import eafp
    from eafp import systems
    from eafp._lib import COEF_BF16, COEF_FP16, COEF_FP32
    fmt = {"fp32": COEF_FP32, "bf16": COEF_BF16, "fp16": COEF_FP16}[args.coef]

    if args.step_only:
        print(json.dumps({"step_only": True, "pp_step": measure_step(eafp)}), flush=True)
        return 0

    if args.tdma_only:
        print(json.dumps({"tdma_only": True, "line_tdma": measure_tdma(eafp)}), flush=True)
        return 0

    if args.assembly_only:
        calibration_copy(CALIB_BYTES)
        print(json.dumps({"assembly_only": True, "pp_assembly": measure_assembly(eafp),
                          "calibration_copy_bytes": CALIB_BYTES}), flush=True)
        return 0

    if args.kernels_only and args.single:
        # the drop-in's two FP64 kernels on the production system
        ps = systems.make_system(2)
        single = eafp.Batch.from_systems([ps])
        if args.tiling:
            shape = [int(v) for v in args.tiling.split(",")]
            single.set_tiling(shape[0], shape[1])
        else:
            single.autotune(eafp.Batch.SOLVER_FP64, reps=5)
        calibration_copy(CALIB_BYTES)
        ms_sp, ms_up = single.time_kernels(eafp.Batch.SOLVER_FP64, reps=20)
        out = {"kernels_only": True, "workload": "single 60x120x84 production system, FP64 drop-in kernels",
               "launch_shape": single.info(), "k64_spmv_avg_us": round(ms_sp * 1e3, 2),
               "k64_update_avg_us": round(ms_up * 1e3, 2), "k64_spmv_algorithmic_bytes": 88 * ps.ncells,
               "k64_update_algorithmic_bytes": 64 * ps.ncells, "calibration_copy_bytes": CALIB_BYTES}
        single.close()
        print(json.dumps(out), flush=True)
        return 0

    t0 = time.time()
    syss = [systems.make_system(5, s) for s in range(B)]
    ncells = sum(p.ncells for p in syss)
    log(f"generated {B} systems 60x120x84 ({ncells} cells) in {time.time() - t0:.1f}s")
    batch = eafp.Batch.from_systems(syss)
    log(f"uploaded; launch shape {batch.info()}")
    tune = None
    msolver = eafp.Batch.SOLVER_MIXED_ZLINE if zline else eafp.Batch.SOLVER_MIXED
    if args.tiling:
        shape = [int(v) for v in args.tiling.split(",")]
        batch.set_tiling(shape[0], shape[1])
        batch.set_columns(shape[2] if len(shape) > 2 else 0)
    elif not args.no_tune:
        tune = batch.autotune(msolver, fmt, reps=5)
        log(f"autotune (mixed {args.coef}): {tune}")
    if args.kernels_only:
        calibration_copy(CALIB_BYTES)
        ms_sp, ms_up = batch.time_kernels(msolver, fmt, reps=20)
        out = {"kernels_only": True, "launch_shape": batch.info(), "coef_format": args.coef,
               "kernels": [spmv_kernel_name(batch.info(), args.coef), update_kernel_name(batch.info(), zline)],
               "k_mspmv_avg_us": round(ms_sp * 1e3, 2), "k_mupdate_avg_us": round(ms_up * 1e3, 2),
               "k_mspmv_algorithmic_bytes": BYTES_MSPMV[args.coef] * ncells,
               "k_mupdate_algorithmic_bytes": BYTES_MUPDATE * ncells,
               "calibration_copy_bytes": CALIB_BYTES}
        batch.close()
        print(json.dumps(out), flush=True)
        return 0

    for w in range(args.warmup):
        st = batch.solve_mixed(tol=1e-8, interval=args.interval, coef_format=fmt, precond=precond)
        log(f"warmup {w}: {st[0]['seconds']:.3f}s")
    device_sync()
    t_start = time.perf_counter()
    stats = []
    for k in range(args.steps):
        stats = batch.solve_mixed(tol=1e-8, interval=args.interval, coef_format=fmt, precond=precond)
    device_sync()
    elapsed = time.perf_counter() - t_start
    value = B * args.steps / elapsed
    ms_per_step = 1e3 * elapsed / args.steps
    conv = all(s["status"] == "converged" for s in stats)
    inner = [s["n_iter"] for s in stats]
    refine = [s["n_refine"] for s in stats]
    res = [s["residual"] for s in stats]
    log(f"timed {args.steps} steps: {ms_per_step:.1f} ms/step, {value:.2f} solves/s, all converged={conv}")
    xs_main = [batch.download(s) for s in range(B)] if not args.no_reference else None

    # the two inner-iteration kernels, timed with HIP events on the library
    # stream; the roofline entry is the one with the longer launch
    ms_sp, ms_up = batch.time_kernels(msolver, fmt, reps=20)
    info = batch.info()
    legs = [("k_mspmv", spmv_kernel_name(info, args.coef), ms_sp, BYTES_MSPMV[args.coef] * ncells),
            ("k_mupdate", update_kernel_name(info, zline), ms_up, BYTES_MUPDATE * ncells)]
    legs.sort(key=lambda l: -l[2])
    rl = []
    for key, name, ms, nbytes in legs:
        ach = nbytes / (ms * 1e-3) / 1e9
        rl.append({"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                   "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": read_traffic(key, workload),
                   "kernel": name, "avg_launch_us": round(ms * 1e3, 2), "algorithmic_bytes_per_launch": nbytes})
    roofline = dict(rl[0])
    roofline["partner"] = rl[1]
    # whole-solve algorithmic bandwidth: inner iterations + refinement steps + setup
    tot_bytes = sum(inner) * (BYTES_MSPMV[args.coef] + BYTES_MUPDATE) * ncells / B
    tot_bytes += sum(refine) * 100.0 * ncells / B + 78.0 * ncells
    solve_gbs = tot_bytes / (elapsed / args.steps) / 1e9

    out = {"metric": METRIC, "value": round(value, 3), "unit": "solves/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 2),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
           "data": "synthetic",
           "config": {"workload": f"batched p' solve: {B} seeded 60x120x84 campaign-variant systems "
                                  "(electrode immersion, slag foam, jittered a_P^u), phi=0, tol 1e-8",
                      "systems": B, "mesh": "60x120x84", "cells": ncells, "tol": 1e-8,
                      "coef_format": args.coef, "refine_interval": args.interval, "precond": args.precond,
                      "launch_shape": batch.info()},
           "roofline": roofline,
           "mixed_batch": {"all_converged": conv, "inner_iters_mean": float(np.mean(inner)),
                           "inner_iters_max": int(max(inner)), "refine_steps_mean": float(np.mean(refine)),
                           "max_true_rel_residual": float(max(res)), "algorithmic_GBps": round(solve_gbs, 1),
                           "autotune": tune}}
    xs_jac = None
    if zline and not args.no_jacobi:
        # the same batch with the Jacobi inner preconditioner, timed the same way (not retuned)
        batch.solve_mixed(tol=1e-8, interval=args.interval, coef_format=fmt)
        device_sync()
        t = time.perf_counter()
        for k in range(args.steps):
            stj = batch.solve_mixed(tol=1e-8, interval=args.interval, coef_format=fmt)
        device_sync()
        tj = (time.perf_counter() - t) / args.steps
        out["mixed_batch"]["jacobi_inner"] = {
            "solves_per_s": round(B / tj, 3), "inner_iters_mean": float(np.mean([x["n_iter"] for x in stj])),
            "all_converged": all(x["status"] == "converged" for x in stj),
            "max_true_rel_residual": float(max(x["residual"] for x in stj))}
        log(f"jacobi inner preconditioner: {B / tj:.2f} solves/s")
        if not args.no_reference:
            xs_jac = [batch.download(s) for s in range(B)]

    # upload from device-resident padded arrays (arrays_on_device = 1, the
    # case of a GPU-resident caller), the singular-component labelling the
    # first mixed solve after an upload runs, and that solve: the rate a
    # caller sees when every solve brings new systems
    up = measure_upload(eafp, batch, syss, fmt, args.interval, precond)
    out["upload"] = up
    out["upload_plus_solve_solves_per_s"] = up["upload_plus_solve_solves_per_s"]
    log(f"upload {up['upload_s']:.3f}s, labelling {up['label_s']:.3f}s, "
        f"upload+label+solve {up['upload_plus_solve_solves_per_s']:.2f} solves/s")

    if not args.no_reference:
        # the plain FP64 GPU CG on the same batch (the drop-in's algorithm, batched):
        # the like-for-like denominator of the mixed speedup; then the same solver to
        # 1e-12 gives the FP64 references of the parity figures
        fb = eafp.Batch.from_systems(syss)
        fb.autotune(eafp.Batch.SOLVER_FP64, reps=3)
        fb.solve_fp64(200000, 1e-8)
        nrep = max(1, min(args.steps, 2))
        t = time.perf_counter()
        for _ in range(nrep):
            sfb = fb.solve_fp64(200000, 1e-8)
        tfb = (time.perf_counter() - t) / nrep
        out["fp64_batch"] = {"workload": f"the same {B} systems, batched FP64 Jacobi-PCG (cg_3d_mpi algorithm)",
                             "solves_per_s": round(B / tfb, 3), "n_iter_mean": float(np.mean([x["n_iter"] for x in sfb])),
                             "all_converged": all(x["status"] == "converged" for x in sfb),
                             "launch_shape": fb.info()}
        out["speedup_mixed_batch_vs_fp64_batch"] = round(value * tfb / B, 3)
        if "jacobi_inner" in out["mixed_batch"]:
            out["speedup_mixed_batch_jacobi_inner_vs_fp64_batch"] = round(
                out["mixed_batch"]["jacobi_inner"]["solves_per_s"] * tfb / B, 3)
        log(f"fp64 batch: {B / tfb:.2f} solves/s")
        refs, stref = fp64_references(eafp, syss, batch=fb)
        fb.close()
        par = {"bar": PARITY_BAR,
               "fp64_reference_converged": all(x["status"] == "converged" for x in stref)}
        par["mixed_batch"] = parity_summary(syss, xs_main, refs)
        if xs_jac is not None:
            par["mixed_batch_jacobi_inner"] = parity_summary(syss, xs_jac, refs)
        out["mixed_batch"]["max_rel_err_vs_fp64_ref"] = par["mixed_batch"]["max_rel_err_p"]
        out["mixed_batch"]["rel_err_checked_systems"] = par["mixed_batch"]["systems_checked"]
        out["parity"] = par
        del xs_main, xs_jac
        log(f"parity of the batch: max rel p' err {par['mixed_batch']['max_rel_err_p']:.2e}, "
            f"max host residual {par['mixed_batch']['max_true_rel_residual_host']:.2e}")

        # the plain FP64 GPU CG (drop-in algorithm) on the single production system
        ps = systems.make_system(2)
        single = eafp.Batch.from_systems([ps])
        single.autotune(eafp.Batch.SOLVER_FP64, reps=5)
        single.solve_fp64(100000, 1e-8)
        t = time.perf_counter()
        reps = 3
        for _ in range(reps):
            s64 = single.solve_fp64(100000, 1e-8)[0]
        t64 = (time.perf_counter() - t) / reps
        x64 = single.download(0)
        ms64 = single.time_kernels(eafp.Batch.SOLVER_FP64, reps=20)
        # the drop-in entry point on host arrays (cg_3d_mpi's call): PCIe copies included
        mesh = eafp.Mesh(*ps.shape)
        arrs = [ps.padded(n) for n in NAMES8]
        phi_h = np.zeros_like(arrs[0], order="F")
        eafp.cg_3d_mpi(*arrs, phi_h, mesh, 100000, 1e-8)
        t = time.perf_counter()
        for _ in range(reps):
            phi_h[...] = 0.0
            eafp.cg_3d_mpi(*arrs, phi_h, mesh, 100000, 1e-8)
        t_drop = (time.perf_counter() - t) / reps
        eafp.release()
        single.autotune(msolver, fmt, reps=5)
        single.solve_mixed(tol=1e-8, interval=args.interval, coef_format=fmt, precond=precond)
        t = time.perf_counter()
        for _ in range(reps):
            smx = single.solve_mixed(tol=1e-8, interval=args.interval, coef_format=fmt, precond=precond)[0]
        tmx = (time.perf_counter() - t) / reps
        xmx = single.download(0)
        ref1, _ = fp64_references(eafp, [ps], batch=single)
        out["fp64_single"] = {"workload": "60x120x84 production system (liquid 7000 below 0.9 m, gas 0.3)",
                              "solves_per_s": round(1.0 / t64, 3), "n_iter": s64["n_iter"],
                              "residual": s64["residual"], "ms_per_iter": round(1e3 * t64 / s64["n_iter"], 4),
                              "kernel_us": {"k64_spmv": round(ms64[0] * 1e3, 2), "k64_update": round(ms64[1] * 1e3, 2)},
                              "algorithmic_GBps": round(BYTES_FP64_ITER * ps.ncells * s64["n_iter"] / t64 / 1e9, 1),
                              "dropin_host_solves_per_s": round(1.0 / t_drop, 3)}
        out["mixed_single"] = {"solves_per_s": round(1.0 / tmx, 3), "inner_iters": smx["n_iter"],
                               "refine_steps": smx["n_refine"], "true_rel_residual": smx["true_rel_l2"],
                               "l1_continuity_residual": smx["true_rel_l1"]}
        par["fp64_single"] = parity_summary([ps], [x64], ref1)
        par["mixed_single"] = parity_summary([ps], [xmx], ref1)
        out["speedup_mixed_batch_vs_fp64"] = round(value * t64, 3)
        out["speedup_mixed_single_vs_fp64"] = round(t64 / tmx, 3)
        single.close()

        if not args.no_cpu_baseline:
            # CPU legs: the oracle restatement of cg_3d_mpi on the production
            # system, on all host threads (OpenMP loops, parallel dots) and on
            # one core (the checker, timed as a baseline only)
            sys.path.insert(0, ROOT)
            import oracle
            oracle.lib()
            oracle.lib_omp()
            threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
            t = time.perf_counter()
            phi_p, res_p, nit_p = oracle.cg_omp(ps, max_iter=100000, tol=1e-8)
            tp = time.perf_counter() - t
            t = time.perf_counter()
            phi_c, res_c, nit_c, _ = oracle.cg(ps, max_iter=100000, tol=1e-8)
            tc = time.perf_counter() - t
            xc = phi_c[2:-2, 2:-2, 2:-2]
            out["cpu_baseline"] = {"value": round(1.0 / tp, 5), "unit": "solves/s", "cores": threads, "kind": "port",
                                   "sample": f"one full FP64 Jacobi-PCG solve (restatement of cg_3d_mpi with OpenMP "
                                             f"loops and parallel dots, {nit_p} iterations to 1e-8) of the "
                                             f"60x120x84 production system, {tp:.2f} s on {threads} host threads",
                                   "one_core": {"value": round(1.0 / tc, 5), "seconds": round(tc, 2),
                                                "n_iter": nit_c}}
            ps1 = systems.make_system(1)
            t = time.perf_counter()
            _, _, nit1p = oracle.cg_omp(ps1, max_iter=100000, tol=1e-8)
            t1p = time.perf_counter() - t
            t = time.perf_counter()
            _, _, nit1, _ = oracle.cg(ps1, max_iter=100000, tol=1e-8)
            t1c = time.perf_counter() - t
            out["cpu_baseline"]["config_12x24x16"] = {"threads_solves_per_s": round(1.0 / t1p, 2),
                                                      "one_core_solves_per_s": round(1.0 / t1c, 2),
                                                      "n_iter": nit1}
            out["mixed_single"]["rel_err_vs_cpu_oracle"] = rel_err(xmx[2:-2, 2:-2, 2:-2], xc, ps)
            out["fp64_single"]["rel_err_vs_cpu_oracle"] = rel_err(x64[2:-2, 2:-2, 2:-2], xc, ps)
            # the GPU FP64 reference (tol 1e-12) against the CPU oracle run to the same tolerance
            phi_r, _, nit_r = oracle.cg_omp(ps, max_iter=400000, tol=1e-12)
            par["fp64_reference_vs_cpu_oracle_1e-12"] = {
                "rel_diff_p": rel_err(ref1[0][2:-2, 2:-2, 2:-2], phi_r[2:-2, 2:-2, 2:-2], ps), "oracle_n_iter": nit_r}
    batch.close()

    if not args.no_extras:
        out["pp_assembly"] = measure_assembly(eafp)
        log("p' assembly done")
        out["pp_step"] = measure_step(eafp)
        log("p' step done")
        out["line_tdma"] = measure_tdma(eafp)
        log("line TDMA done")
        # the other workloads of the metric's table, mixed against FP64 on the same GPU
        out["other_workloads"] = []
        syss3 = [systems.make_system(3, s) for s in range(256)]
        out["other_workloads"].append(measure_workload(
            eafp, syss3, fmt, args.coef, args.interval,
            "256 ragged systems U{18..35}x{36..66}x{24..50}, liquid height and gas density jittered", precond))
        del syss3
        log("workload 256 ragged systems done")
        ps4 = systems.make_system(4, 0)
        out["other_workloads"].append(measure_workload(
            eafp, [ps4], COEF_BF16, "bf16", args.interval,
            "one 120x240x168 system, scrap band 0.9<z<2.5 m with face coefficients x1e-2(1-alpha_s), BF16 weights",
            precond))
        del ps4
        log("workload 120x240x168 done")
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
