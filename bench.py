#!/usr/bin/env python3
"""Benchmark: nonlinear heat on a 2048 x 2048 grid, Newton to 1e-8 relative residual.

One step = one implicit time step (one newtonSolve: residual + Jacobian assembly by
the device AD engine, Krylov solve, update) of u_t = div(k(u) grad u),
k(u) = 0.1 (1 + u^2), dt = 1, on device-resident state.  The headline run uses the
mixed-precision solver (FP32 BiCGStab + multicolour ILU(0), FP64 residual, Jacobian,
defect and convergence test); the plain FP64 solver runs on the same input for
comparison.  Metric: cell-Newton-iterations per second = cells x Newton iterations /
wall seconds over the timed steps.

roofline: the Krylov SpMV (dominant kernel), algorithmic bytes per launch
8 nnz + 12 n + 4 (FP32 values, int32 indices, x read once, y written once),
timed live with HIP events on the runtime stream over the timed steps.
cpu_baseline: the CPU oracle (serial C++ restatement of the reference serial
backend) timed for one Newton iteration of the same workload on one host core.

  python bench.py [--steps K] [--warmup W] [--size N] [--no-cpu-baseline]
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from equelle_amd.capi import CApi, dptr, iptr  # noqa: E402
from equelle_amd.runtime import load_library  # noqa: E402
from equelle_amd.workloads import NONLINEAR_HEAT, gaussian_bumps, heat_params  # noqa: E402

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def run_heat(lib, params, u0, dt, steps, warmup):
    n = u0.size
    sim = C.c_void_p()
    e, ed = np.zeros(1, dtype=np.int32), np.zeros(1)
    t0 = time.perf_counter()
    lib.heat_create(params.encode(), dptr(u0), n, iptr(e), dptr(ed), 0, C.byref(sim))
    setup_s = time.perf_counter() - t0
    stats = np.zeros(3, dtype=np.int32)
    norms = np.zeros(2)
    try:
        for w in range(warmup):
            t = time.perf_counter()
            lib.heat_step(sim, dt, iptr(stats), dptr(norms))
            log(f"  warmup step {w}: newton={stats[0]} krylov={stats[1]} stop={stats[2]} "
                f"{time.perf_counter() - t:.3f}s")
        lib.device_sync()
        lib.timing(2)
        lib.timing(1)
        newton = krylov = 0
        per_step = []
        t0 = time.perf_counter()
        for s in range(steps):
            t = time.perf_counter()
            lib.heat_step(sim, dt, iptr(stats), dptr(norms))
            per_step.append(time.perf_counter() - t)
            newton += int(stats[0])
            krylov += int(stats[1])
            if stats[2] == 0:
                raise RuntimeError(f"Newton did not converge in step {s}: ||F|| {norms[0]:.3e} -> {norms[1]:.3e}")
            log(f"  step {s}: newton={stats[0]} krylov={stats[1]} rel={norms[1] / norms[0]:.2e} {per_step[-1]:.3f}s")
        lib.device_sync()
        secs = time.perf_counter() - t0
        lib.timing(0)
        spmv = {}
        for name in ("spmv_fp32", "spmv_fp64"):
            cnt, tot = C.c_int(), C.c_double()
            lib.timing_query(name.encode(), C.byref(cnt), C.byref(tot))
            if cnt.value:
                spmv[name] = (cnt.value, tot.value / cnt.value)
        u = np.zeros(n)
        lib.heat_state(sim, dptr(u))
    finally:
        lib.heat_destroy(sim)
    return dict(seconds=secs, newton=newton, krylov=krylov, per_step=per_step, spmv=spmv, setup_s=setup_s,
                u=u)


def cpu_baseline(n_side, u0, dt):
    """Serial oracle: one Newton iteration (AD assembly + BiCGStab/ILU(0) to the same
    relative tolerance) on the same grid and state."""
    path = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(path):
        return None
    api = CApi(path, "eqo_")
    params = (f"grid_dim=2 nx={n_side} ny={n_side} k={NONLINEAR_HEAT['k']} cv=1 k_model=quadratic "
              "oracle_linear_solver=bicgstab oracle_solver_tol=1e-6").encode()
    secs, lin = C.c_double(), C.c_int()
    api.heat_newton_sample(params, dptr(u0), u0.size, 1, dt, C.byref(secs), C.byref(lin))
    return dict(value=u0.size / secs.value, unit="cell-Newton-iters/s", cores=1, kind="port",
                sample=f"1 Newton iteration (AD residual+Jacobian, BiCGStab/ILU(0) to 1e-6, {lin.value} "
                       f"iterations) of the {n_side}x{n_side} nonlinear-heat step, serial C++ oracle: "
                       f"{secs.value:.2f} s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=NONLINEAR_HEAT["nx"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP64 comparison run")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    lib = load_library()
    nx = ny = args.size
    n = nx * ny
    nif = (nx - 1) * ny + nx * (ny - 1)
    nnz = n + 2 * nif
    dt = NONLINEAR_HEAT["dt"]
    u0 = gaussian_bumps(nx, ny, seed=2048)
    runs = {}
    for prec in (["mixed"] if args.no_plain else ["mixed", "fp64"]):
        log(f"[bench] nonlinear heat {nx}x{ny}, precision={prec}")
        params = heat_params(nx, ny, NONLINEAR_HEAT["k"], precision=prec, precond="ilu0",
                             rel_res_tol=NONLINEAR_HEAT["rel_res_tol"], solver_tol=1e-6)
        runs[prec] = run_heat(lib, params, u0, dt, args.steps, args.warmup)
    head = runs["mixed"]
    value = n * head["newton"] / head["seconds"]
    cnt, avg_ms = head["spmv"].get("spmv_fp32", (0, float("nan")))
    spmv_bytes = 8 * nnz + 12 * n + 4
    achieved = spmv_bytes / (avg_ms * 1e-3) / 1e9
    out = {
        "metric": "cell-Newton-iters/s",
        "value": value,
        "unit": "cell-Newton-iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * head["seconds"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 residual/Jacobian/refinement, f32 Krylov",
        "data": "synthetic",
        "config": {"workload": f"nonlinear heat {nx}x{ny}, k(u)=0.1(1+u^2), dt=1, Newton to 1e-8 relative "
                               "residual, mixed precision (FP32 BiCGStab+ILU0, FP64 refinement)",
                   "cells": n, "jacobian_nnz": nnz, "engine": "device AD engine (generic program path)"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_sell_spmv<float>", "launches": cnt, "avg_us": 1e3 * avg_ms,
                     "algorithmic_bytes": spmv_bytes},
        "newton_iterations": head["newton"],
        "krylov_iterations": head["krylov"],
        "cell_steps_per_s": n * args.steps / head["seconds"],
    }
    if "fp64" in runs:
        p = runs["fp64"]
        pv = n * p["newton"] / p["seconds"]
        c64, a64 = p["spmv"].get("spmv_fp64", (0, float("nan")))
        b64 = 12 * nnz + 20 * n + 4
        out["plain_fp64"] = {"value": pv, "ms_per_step": 1e3 * p["seconds"] / args.steps,
                             "newton_iterations": p["newton"], "krylov_iterations": p["krylov"],
                             "spmv_avg_us": 1e3 * a64, "spmv_GBs": b64 / (a64 * 1e-3) / 1e9}
        out["mixed_over_plain"] = value / pv
        rel = np.linalg.norm(head["u"] - p["u"]) / np.linalg.norm(p["u"])
        out["mixed_vs_fp64_rel_error"] = float(rel)
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline (serial oracle, one Newton iteration)")
        out["cpu_baseline"] = cpu_baseline(nx, u0, dt)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
