"""p' CG solves/s at 60x120x84 to 1e-8 relative residual: mixed precision
(batched) vs the plain FP64 GPU CG, with the roofline of the dominant kernel
and the CPU oracle beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--systems B]

One step = one batched mixed-precision solve (FP32 inner CG on compressed
face weights + FP64 refinement, setup pass included) of B seeded 60x120x84
campaign-variant systems already resident in HBM, from phi = 0 to a true
relative residual of 1e-8 on every system.  `value` = B*K / elapsed seconds.

Beside it, on the same GPU: the plain FP64 Jacobi-PCG (the cg_3d_mpi
algorithm, the drop-in path) on the single 60x120x84 production system, the
mixed solver on that same single system, the relative p' error of the mixed
results against FP64 reference solutions, and the CPU restatement of
cg_3d_mpi (1 core) on the production system.
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
BYTES_MSPMV = {"fp32": 32, "bf16": 26, "fp16": 30}   # z, p_old, 3 weights (+scale), e; p, q
BYTES_MUPDATE = 32                                   # d, r, p, q, 1/D; d, r, z
BYTES_FP64_ITER = 152                                # k64_spmv 88 + k64_update 64


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


def read_traffic(kernel, workload):
    """HBM bytes per launch measured with rocprofv3 counters (profiles/), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)
        return rec.get(workload, {}).get(kernel, {}).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


def rel_err(x, xr, ps):
    from eafp.systems import project_zero_mean
    px = project_zero_mean(x, ps.labels, ps.ncomp)
    pr = project_zero_mean(xr, ps.labels, ps.ncomp)
    return float(np.linalg.norm(px - pr) / max(np.linalg.norm(pr), 1e-300))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--systems", type=int, default=64)
    ap.add_argument("--coef", choices=("fp32", "bf16", "fp16"), default="fp32")
    ap.add_argument("--interval", type=int, default=32)
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-reference", action="store_true", help="skip the FP64 legs and error checks")
    ap.add_argument("--tiling", default=None, help="pin the launch shape 'cells_per_thread,blocks' (no autotune)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="only the roofline leg: a 1 GiB device copy (counter calibration) and the timed "
                         "k_mspmv/k_mupdate launches, as measured for the bench line")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (--gpus {args.gpus} requested); "
              "the batched p' solve is sharded by system across GPUs in a multi-GPU build, which this is not.",
              file=sys.stderr)
        return 2

    import eafp
    from eafp import systems
    from eafp._lib import COEF_BF16, COEF_FP16, COEF_FP32
    fmt = {"fp32": COEF_FP32, "bf16": COEF_BF16, "fp16": COEF_FP16}[args.coef]
    B = args.systems

    t0 = time.time()
    syss = [systems.make_system(5, s) for s in range(B)]
    ncells = sum(p.ncells for p in syss)
    log(f"generated {B} systems 60x120x84 ({ncells} cells) in {time.time() - t0:.1f}s")
    batch = eafp.Batch.from_systems(syss)
    log(f"uploaded; launch shape {batch.info()}")
    tune = None
    if args.tiling:
        cpt, blocks = (int(v) for v in args.tiling.split(","))
        batch.set_tiling(cpt, blocks)
    elif not args.no_tune:
        tune = batch.autotune(eafp.Batch.SOLVER_MIXED, fmt, reps=5)
        log(f"autotune (mixed {args.coef}): {tune}")
    if args.kernels_only:
        calibration_copy(1 << 30)
        ms_sp, ms_up = batch.time_kernels(eafp.Batch.SOLVER_MIXED, fmt, reps=20)
        out = {"kernels_only": True, "launch_shape": batch.info(), "coef_format": args.coef,
               "k_mspmv_avg_us": round(ms_sp * 1e3, 2), "k_mupdate_avg_us": round(ms_up * 1e3, 2),
               "k_mspmv_algorithmic_bytes": BYTES_MSPMV[args.coef] * ncells,
               "k_mupdate_algorithmic_bytes": BYTES_MUPDATE * ncells,
               "calibration_copy_bytes": 1 << 30}
        batch.close()
        print(json.dumps(out), flush=True)
        return 0

    for w in range(args.warmup):
        st = batch.solve_mixed(tol=1e-8, interval=args.interval, coef_format=fmt)
        log(f"warmup {w}: {st[0]['seconds']:.3f}s")
    device_sync()
    t_start = time.perf_counter()
    stats = []
    for k in range(args.steps):
        stats = batch.solve_mixed(tol=1e-8, interval=args.interval, coef_format=fmt)
    device_sync()
    elapsed = time.perf_counter() - t_start
    value = B * args.steps / elapsed
    ms_per_step = 1e3 * elapsed / args.steps
    conv = all(s["status"] == "converged" for s in stats)
    inner = [s["n_iter"] for s in stats]
    refine = [s["n_refine"] for s in stats]
    res = [s["residual"] for s in stats]
    log(f"timed {args.steps} steps: {ms_per_step:.1f} ms/step, {value:.2f} solves/s, all converged={conv}")

    # dominant kernel (k_mspmv) and its partner, timed with HIP events on the library stream
    ms_sp, ms_up = batch.time_kernels(eafp.Batch.SOLVER_MIXED, fmt, reps=20)
    b_sp = BYTES_MSPMV[args.coef] * ncells
    b_up = BYTES_MUPDATE * ncells
    ach = b_sp / (ms_sp * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": read_traffic("k_mspmv", "batch64_60x120x84"),
                "kernel": f"k_mspmv<{args.coef}>", "avg_launch_us": round(ms_sp * 1e3, 2),
                "algorithmic_bytes_per_launch": b_sp,
                "partner": {"kernel": "k_mupdate", "avg_launch_us": round(ms_up * 1e3, 2),
                            "achieved": round(b_up / (ms_up * 1e-3) / 1e9, 1)}}
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
                      "coef_format": args.coef, "refine_interval": args.interval,
                      "launch_shape": batch.info()},
           "roofline": roofline,
           "mixed_batch": {"all_converged": conv, "inner_iters_mean": float(np.mean(inner)),
                           "inner_iters_max": int(max(inner)), "refine_steps_mean": float(np.mean(refine)),
                           "max_true_rel_residual": float(max(res)), "algorithmic_GBps": round(solve_gbs, 1),
                           "autotune": tune}}

    if not args.no_reference:
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
        x64 = single.download(0)[2:-2, 2:-2, 2:-2].copy()
        ms64 = single.time_kernels(eafp.Batch.SOLVER_FP64, reps=20)
        single.autotune(eafp.Batch.SOLVER_MIXED, fmt, reps=5)
        single.solve_mixed(tol=1e-8, interval=args.interval, coef_format=fmt)
        t = time.perf_counter()
        for _ in range(reps):
            smx = single.solve_mixed(tol=1e-8, interval=args.interval, coef_format=fmt)[0]
        tmx = (time.perf_counter() - t) / reps
        xmx = single.download(0)[2:-2, 2:-2, 2:-2].copy()
        out["fp64_single"] = {"workload": "60x120x84 production system (liquid 7000 below 0.9 m, gas 0.3)",
                              "solves_per_s": round(1.0 / t64, 3), "n_iter": s64["n_iter"],
                              "residual": s64["residual"], "ms_per_iter": round(1e3 * t64 / s64["n_iter"], 4),
                              "kernel_us": {"k64_spmv": round(ms64[0] * 1e3, 2), "k64_update": round(ms64[1] * 1e3, 2)},
                              "algorithmic_GBps": round(BYTES_FP64_ITER * ps.ncells * s64["n_iter"] / t64 / 1e9, 1)}
        out["mixed_single"] = {"solves_per_s": round(1.0 / tmx, 3), "inner_iters": smx["n_iter"],
                               "refine_steps": smx["n_refine"], "true_rel_residual": smx["true_rel_l2"],
                               "l1_continuity_residual": smx["true_rel_l1"]}
        out["speedup_mixed_batch_vs_fp64"] = round(value * t64, 3)
        out["speedup_mixed_single_vs_fp64"] = round(t64 / tmx, 3)
        single.close()

        # relative p' error of the batched mixed results against FP64 reference
        # solutions of the same systems (GPU FP64 CG at tol 1e-11, the drop-in
        # algorithm whose parity with the CPU oracle the test suite checks)
        nchk = min(B, 8)
        ref = eafp.Batch.from_systems(syss[:nchk])
        ref.solve_fp64(200000, 1e-11)
        errs = [rel_err(batch.download(s)[2:-2, 2:-2, 2:-2], ref.download(s)[2:-2, 2:-2, 2:-2], syss[s])
                for s in range(nchk)]
        ref.close()
        out["mixed_batch"]["max_rel_err_vs_fp64_ref"] = max(errs)
        out["mixed_batch"]["rel_err_checked_systems"] = nchk

        if not args.no_cpu_baseline:
            # CPU leg: the oracle restatement of cg_3d_mpi, 1 core, on the
            # production system (the checker, timed as the baseline only)
            sys.path.insert(0, ROOT)
            import oracle
            oracle.lib()
            t = time.perf_counter()
            phi_c, res_c, nit_c, _ = oracle.cg(ps, max_iter=100000, tol=1e-8)
            tc = time.perf_counter() - t
            xc = phi_c[2:-2, 2:-2, 2:-2]
            out["cpu_baseline"] = {"value": round(1.0 / tc, 5), "unit": "solves/s", "cores": 1, "kind": "port",
                                   "sample": f"one full FP64 Jacobi-PCG solve (oracle restatement of cg_3d_mpi, "
                                             f"{nit_c} iterations to 1e-8) of the 60x120x84 production system, "
                                             f"{tc:.1f} s on 1 host core"}
            out["mixed_single"]["rel_err_vs_cpu_oracle"] = rel_err(xmx, xc, ps)
            out["fp64_single"]["rel_err_vs_cpu_oracle"] = rel_err(x64, xc, ps)
    batch.close()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
