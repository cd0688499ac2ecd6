#!/usr/bin/env python3
"""Benchmark: nonlinear heat on a 2048 x 2048 grid, Newton to 1e-8 relative residual.

One step = one implicit time step (one Newton solve: residual + Jacobian assembly,
Krylov solve, update) of u_t = div(k(u) grad u), k(u) = 0.1 (1 + u^2), dt = 1,
zero-flux boundary, on device-resident state.  Metric: cell-Newton-iterations per
second = cells x Newton iterations / wall seconds over the timed steps.

Runs on the same seeded input (16 Gaussian bumps + noise):
  headline     fused engine, mixed precision: FP32 BiCGStab on the red-black reduced
               (Schur complement) system, preconditioned by its diagonal (= red-black
               ILU(0)), inner tolerance 1e-6; FP64 residual, Jacobian, defect (iterative
               refinement to 1e-6) and Newton test
  plain_fp64   fused engine, the same Krylov method in FP64
  generic_fp64 the drop-in path: the same program through EquelleRuntimeCUDA's
               device AD engine (CSR Jacobian, SELL-64 SpMV, multicolour ILU(0))
kernels: every kernel of the headline's Newton and Krylov loop, timed on the final state
of the timed run: 50 launches each between their own HIP start/stop events on the
runtime stream (avg_us), and 50 more back to back (back_to_back_us); algorithmic bytes
per launch as stated per kernel.  roofline: the kernel with the largest estimated time
per step (avg_us x launches per step from the iteration counts); roofline.traffic: its HBM
bytes per launch measured in the same run (rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE over
child runs of this script, scaled by the counters of a 256 MiB device-to-device copy).
cpu_baseline: the CPU oracle (serial C++ restatement of the reference serial backend)
timed for one Newton iteration of the same workload on one host core.
Every leg runs its W warmup steps and is then reset to its initial state, so the K timed
steps are the workload's steps 1..K.

The default run also carries compact summaries of the other configured workloads:
  batched      512 nonlinear-heat instances (sides 32..512) in one mixed-precision batch
               against the plain FP64 path one instance at a time
  twophase     the 3D two-phase workload (128 x 128 x 64), all 30 configured steps: the
               fused engine (FP64 and mixed) and the configured method (FP32 GMRES(30) +
               Jacobi with FP64 refinement on the drop-in AD engine)
  serving      the request loop at two Poisson arrival rates, p50/p99 per request kind

  python bench.py [--steps K] [--warmup W] [--size N] [--no-cpu-baseline] [--quick] [--no-extras]
                  [--no-traffic]
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
FORCING = ""   # "forcing=<x>" from --forcing (fused runs)
COPY_BYTES = 256 << 20   # the counter calibration copies 256 MiB device to device per launch


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def timing_avg(lib, name):
    cnt, tot = C.c_int(), C.c_double()
    lib.timing_query(name.encode(), C.byref(cnt), C.byref(tot))
    return cnt.value, (tot.value / cnt.value if cnt.value else float("nan"))


def run_fused(lib, precision, u0, nx, ny, dt, steps, warmup, precond="ilu0", graph=True, inner_tol=1e-6, extra=""):
    params = (f"precision={precision} preconditioner={precond} rel_res_tol={NONLINEAR_HEAT['rel_res_tol']} "
              f"abs_res_tol=0 max_iter=20 solver_tol=1e-6 inner_tol={inner_tol} graph={int(graph)} {FORCING} "
              f"{extra}").encode()
    b = C.c_void_p()
    lib.batch_create(params, C.byref(b))
    slot = C.c_int()
    st = np.zeros(5, dtype=np.int32)
    nrm = np.zeros(2)
    try:
        lib.batch_add(b, nx, ny, NONLINEAR_HEAT["k"], 1.0, dptr(u0), C.byref(slot))
        slots = np.array([slot.value], dtype=np.int32)
        for w in range(warmup):
            t = time.perf_counter()
            lib.batch_step(b, iptr(slots), 1, dt)
            lib.batch_stats(b, slot.value, iptr(st), dptr(nrm))
            log(f"  warmup {w}: newton={st[0]} krylov={st[1]} refine={st[2]} {time.perf_counter() - t:.4f}s")
        lib.batch_set_state(b, slot.value, dptr(u0))   # the timed steps are steps 1..K from u0
        lib.device_sync()
        newton = krylov = refine = 0
        lines = []   # printed after the timed steps
        t0 = time.perf_counter()
        for s in range(steps):
            t = time.perf_counter()
            lib.batch_step(b, iptr(slots), 1, dt)
            lib.batch_stats(b, slot.value, iptr(st), dptr(nrm))
            if st[3] == 0 or st[4]:
                raise RuntimeError(f"step {s} did not converge: stats {st.tolist()} norms {nrm.tolist()}")
            newton += int(st[0])
            krylov += int(st[1])
            refine += int(st[2])
            lines.append((s, int(st[0]), int(st[1]), int(st[2]), nrm[1] / nrm[0], time.perf_counter() - t))
        lib.device_sync()
        secs = time.perf_counter() - t0
        for s, nw, kr, rf, rel, dtm in lines:
            log(f"  step {s}: newton={nw} krylov={kr} refine={rf} rel={rel:.2e} {dtm:.4f}s")
        # kernel durations on the final state, every instance active: 50 launches of each
        # stencil kernel over the whole grid, each between its own HIP start/stop events on
        # the runtime stream (avg), then 50 more back to back (dispatch gaps included)
        shape = np.zeros(3, dtype=np.int32)
        lib.batch_shape(b, iptr(shape))
        cls = np.zeros(5, dtype=np.int32)
        lib.batch_shape_classes(b, iptr(cls))
        kern = {"shape": (int(shape[0]), int(shape[1])),
                "cpt": dict(zip(("assemble", "defect", "spmv", "precond", "vector"), map(int, cls)))}
        # the reduced form with nx % 8 == 0 runs its Krylov vectors four ranks per lane, whose
        # iterations end in one fused update (b31) instead of b1 and b3
        vec = precond == "ilu0" and nx % 8 == 0
        kern["tiles"] = vec and nx % 512 == 0
        kern["vec"] = vec
        names = (("operator",) if vec else ("operator_red", "operator_black") if precond == "ilu0" else ("spmv",)) + \
            ("assemble",) + (("defect",) if precision == "mixed" else ()) + \
            (("b2", "b31", "setup") if vec else ("b1", "b2", "b3", "load_rhs"))
        for name in names:
            ms, ev = C.c_double(), C.c_double()
            lib.batch_kernel_time(b, name.encode(), 50, C.byref(ms), C.byref(ev))
            kern[name] = (50, ms.value)
            kern[name + "_event"] = ev.value
        u = np.zeros(nx * ny)
        lib.batch_state(b, slot.value, dptr(u))
    finally:
        lib.batch_destroy(b)
    return dict(seconds=secs, newton=newton, krylov=krylov, refine=refine, kern=kern, u=u)


def run_generic(lib, u0, nx, ny, dt, steps, warmup, precision="fp64", extra="", eval_reps=0):
    params = heat_params(nx, ny, NONLINEAR_HEAT["k"], precision=precision, precond="ilu0",
                         rel_res_tol=NONLINEAR_HEAT["rel_res_tol"], solver_tol=1e-6, extra=extra)
    sim = C.c_void_p()
    e, ed = np.zeros(1, dtype=np.int32), np.zeros(1)
    t0 = time.perf_counter()
    lib.heat_create(params.encode(), dptr(u0), u0.size, iptr(e), dptr(ed), 0, C.byref(sim))
    stats = np.zeros(3, dtype=np.int32)
    first = None
    try:
        for _ in range(warmup):
            lib.heat_step(sim, dt, iptr(stats), None)
            if first is None:   # set-up and the first step: grid, symbolic passes, tuner
                lib.device_sync()
                first = time.perf_counter() - t0
        lib.heat_set_state(sim, dptr(u0))   # timed steps are steps 1..K from u0
        lib.device_sync()
        newton = krylov = 0
        t0 = time.perf_counter()
        for _ in range(steps):
            lib.heat_step(sim, dt, iptr(stats), None)
            newton += int(stats[0])
            krylov += int(stats[1])
        lib.device_sync()
        secs = time.perf_counter() - t0
        u = np.zeros(nx * ny)
        lib.heat_state(sim, dptr(u))
        assembly = None
        if eval_reps:
            ms, info = C.c_double(), np.zeros(3, dtype=np.int32)
            lib.heat_eval_time(sim, dt, eval_reps, C.byref(ms), iptr(info))
            assembly = assembly_roofline(nx, ny, int(info[1]), ms.value, eval_reps)
            assembly["evaluation"] = "one fused kernel (recorded residual program)" if info[2] else "op by op"
    finally:
        lib.heat_destroy(sim)
    return dict(seconds=secs, newton=newton, krylov=krylov, u=u, first_step_s=first, assembly=assembly)


def assembly_roofline(nx, ny, nnz, ms, reps):
    """One residual + Jacobian evaluation of the drop-in program against the assembly floor:
    8 nnz (Jacobian values written) + 8 n (F written) + 16 n (u, u0 read) + 8 nif (the
    per-face transmissibility read) bytes."""
    n, nif = nx * ny, (nx - 1) * ny + nx * (ny - 1)
    floor = 8 * nnz + 8 * n + 16 * n + 8 * nif
    gbs = floor / (ms * 1e-3) / 1e9
    return {"ms_per_eval": ms, "evaluations": reps, "jacobian_nnz": nnz, "algorithmic_bytes": floor,
            "bytes_formula": "8 nnz + 24 n + 8 nif", "GBs": gbs, "frac": gbs / HBM_PEAK_GBS,
            "measured": "host clock over the evaluations (kernels and host work of the AD engine), "
                        "after one untimed evaluation"}


def bench_generic(lib, args):
    """The drop-in path alone: the nonlinear heat program through EquelleRuntimeCUDA's
    device AD engine (one newtonSolve per step, BiCGStab + ILU(0)), FP64 and mixed
    precision, with the Jacobian solved on the stencil kernels (its pattern is the grid's
    5-point stencil) and, for comparison, on the CSR / SELL-64 kernels (stencil_solver=0)."""
    nx = ny = args.size
    n = nx * ny
    u0 = gaussian_bumps(nx, ny, seed=2048)
    legs = {}
    ref = None
    assembly = None
    assembly_op = None
    for name, prec, extra in (("fp64_stencil", "fp64", ""), ("mixed_stencil", "mixed", "inner_tol=1e-6"),
                              ("fp64_stencil_op_by_op", "fp64", "fuse_residual=0"),
                              ("fp64_csr", "fp64", "stencil_solver=0"),
                              ("mixed_csr", "mixed", "stencil_solver=0 inner_tol=1e-6")):
        if args.quick and (name.endswith("_csr") or name.endswith("op_by_op")):
            continue
        log(f"[bench] generic drop-in path {name}")
        g = run_generic(lib, u0, nx, ny, NONLINEAR_HEAT["dt"], args.steps, args.warmup, prec, extra,
                        eval_reps=10 if name in ("fp64_stencil", "fp64_stencil_op_by_op") else 0)
        if ref is None:
            ref = g["u"]
        if g["assembly"] is not None:
            if name == "fp64_stencil":
                assembly = g["assembly"]
            else:
                assembly_op = g["assembly"]
        legs[name] = {"value": n * g["newton"] / g["seconds"], "ms_per_step": 1e3 * g["seconds"] / args.steps,
                      "newton_iterations": g["newton"], "krylov_iterations": g["krylov"],
                      "setup_and_first_step_s": g["first_step_s"],
                      "rel_error_vs_fp64_stencil": float(np.linalg.norm(g["u"] - ref) / np.linalg.norm(ref))}
    head = legs["fp64_stencil"]
    out = {"metric": "cell-Newton-iters/s", "value": head["value"], "unit": "cell-Newton-iters/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": head["ms_per_step"], "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"nonlinear heat {nx}x{ny} through the drop-in EquelleRuntimeCUDA AD engine, "
                                  "BiCGStab + ILU(0), Newton to 1e-8 relative residual", "cells": n},
           "legs": legs, "assembly_roofline": assembly, "assembly_roofline_op_by_op": assembly_op}
    print(json.dumps(out), flush=True)


def batched_runs(lib, args, steps, warmup, seq_steps, generic=True):
    """512 independent nonlinear-heat instances (sides log-uniform 32..512, k0 ~ U[0.05,0.5]).

    batched_mixed         one fused batch, mixed precision, all instances in every launch
                          (`steps` timed steps)
    sequential_fp64_fused the plain FP64 path of the same engine, one instance at a time
                          (`seq_steps` timed steps: a rate)
    sequential_fp64_generic  the drop-in AD engine (FP64), one instance at a time, on every
                          8th instance (a rate, so the subsample keeps the size mix)
    Every leg runs its warmup steps, then resets each instance to its initial field, so the
    timed steps are the workload's first steps."""
    from equelle_amd.workloads import batched_instances
    inst = batched_instances(args.instances, seed=1234)
    cells = np.array([s * s for s, _, _ in inst], dtype=np.int64)
    dt = 1.0
    common = f"rel_res_tol={NONLINEAR_HEAT['rel_res_tol']} abs_res_tol=0 max_iter=20 solver_tol=1e-6"

    def run(prec, batched, nsteps):
        b = C.c_void_p()
        lib.batch_create(f"precision={prec} preconditioner={args.precond} {common}".encode(), C.byref(b))
        try:
            slots = []
            for side, k0, u0 in inst:
                sl = C.c_int()
                lib.batch_add(b, side, side, k0, 1.0, dptr(u0), C.byref(sl))
                slots.append(sl.value)
            slots = np.array(slots, dtype=np.int32)
            groups = [slots] if batched else [slots[q:q + 1] for q in range(len(slots))]
            st = np.zeros((len(slots), 5), dtype=np.int32)

            def one_step():
                for g in groups:
                    lib.batch_step(b, iptr(g), len(g), dt)
                lib.batch_stats_many(b, iptr(slots), len(slots), iptr(st))   # host-side counters
                bad = np.nonzero((st[:, 3] == 0) | (st[:, 4] != 0))[0]
                if bad.size:
                    raise RuntimeError(f"instance {bad[0]} did not converge: {st[bad[0]].tolist()}")
                return int(np.dot(cells, st[:, 0].astype(np.int64)))

            for _ in range(warmup):
                one_step()
            for (_, _, u0), sl in zip(inst, slots):
                lib.batch_set_state(b, int(sl), dptr(u0))
            lib.device_sync()
            work = 0
            t0 = time.perf_counter()
            for _ in range(nsteps):
                work += one_step()
            lib.device_sync()
            secs = time.perf_counter() - t0
            u = [np.zeros(s * s) for s, _, _ in inst]
            for q, sl in enumerate(slots):
                lib.batch_state(b, int(sl), dptr(u[q]))
            shape = np.zeros(3, dtype=np.int32)
            lib.batch_shape(b, iptr(shape))
            return dict(value=work / secs, ms_per_step=1e3 * secs / nsteps, steps=nsteps, newton_cells=work,
                        launch_shape={"cells_per_thread": int(shape[0]), "xcd_block_order": int(shape[1]),
                                      "signatures_tuned": int(shape[2])}), u
        finally:
            lib.batch_destroy(b)

    log(f"[bench] batched: {len(inst)} instances, {cells.sum()} cells, mixed precision, one batch")
    res = {}
    res["batched_mixed"], ub = run("mixed", True, steps)
    log("[bench] batched: plain FP64, one instance at a time (fused engine)")
    res["sequential_fp64_fused"], us = run("fp64", False, seq_steps)
    if seq_steps == steps:
        err = max(float(np.linalg.norm(a - c) / np.linalg.norm(c)) for a, c in zip(ub, us))
    else:
        err = None
    if generic:
        log("[bench] batched: plain FP64, one instance at a time (generic drop-in AD engine), every 8th instance")
        e, ed = np.zeros(1, dtype=np.int32), np.zeros(1)
        stats = np.zeros(3, dtype=np.int32)
        work, secs = 0, 0.0
        for side, k0, u0 in inst[::8]:
            sim = C.c_void_p()
            params = heat_params(side, side, k0, precision="fp64", precond="ilu0",
                                 rel_res_tol=NONLINEAR_HEAT["rel_res_tol"], solver_tol=1e-6)
            lib.heat_create(params.encode(), dptr(u0), side * side, iptr(e), dptr(ed), 0, C.byref(sim))
            try:
                lib.heat_step(sim, dt, iptr(stats), None)     # first call builds the symbolic cache
                lib.heat_set_state(sim, dptr(u0))
                lib.device_sync()
                t0 = time.perf_counter()
                lib.heat_step(sim, dt, iptr(stats), None)
                lib.device_sync()
                secs += time.perf_counter() - t0
                work += side * side * int(stats[0])
            finally:
                lib.heat_destroy(sim)
        res["sequential_fp64_generic"] = dict(value=work / secs, instances=len(inst[::8]), steps=1)
    head = res["batched_mixed"]
    out = {"value": head["value"], "unit": "cell-Newton-iters/s", "ms_per_step": head["ms_per_step"],
           "instances": len(inst), "cells": int(cells.sum()), "runs": res,
           "batched_over_sequential_fp64_fused": head["value"] / res["sequential_fp64_fused"]["value"]}
    if generic:
        out["batched_over_sequential_fp64_generic"] = head["value"] / res["sequential_fp64_generic"]["value"]
    if err is not None:
        out["mixed_vs_fp64_max_rel_error"] = err
    return out


def bench_batched(lib, args):
    """The batched workload on its own: one JSON line (value = batched mixed precision)."""
    r = batched_runs(lib, args, args.steps, args.warmup, args.steps)
    out = {"metric": "cell-Newton-iters/s", "value": r["value"], "unit": "cell-Newton-iters/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": r["ms_per_step"],
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64 residual/Jacobian/refinement, f32 Krylov", "data": "synthetic",
           "config": {"workload": f"{r['instances']} nonlinear-heat instances in one batch, sides log-uniform "
                                  "32..512, k0 ~ U[0.05,0.5], dt=1, Newton to 1e-8 relative residual each",
                      "cells": r["cells"], "precond": args.precond}}
    out.update({k: v for k, v in r.items() if k not in ("value", "unit", "ms_per_step")})
    print(json.dumps(out), flush=True)


def twophase_runs(lib, args, steps, warmup, generic_legs):
    """3D two-phase transport (twophase.equelle, Corey 2, viscosity ratio 5): per step a
    pressure Newton solve and a saturation Newton solve, each to 1e-8 relative residual.
    Work = cells x (pressure + transport Newton iterations).

    fused_fp64    fused stencil engine, FP64: the pressure stage by CG with the aggregation
                  multigrid V-cycle, the saturation stage by BiCGStab with the red-black ILU(0)
    fused_mixed   the same with FP32 Krylov and FP64 refinement; a step that does not
                  converge is recorded under runs.fused_mixed.failed instead of ending the run
    generic_*     the drop-in path: the printer-style program on the device AD engine;
                  generic_legs = [(name, solver settings, steps)], e.g. the configured
                  method FP32 GMRES(30) + Jacobi with FP64 refinement
    Every leg runs its warmup steps and is then reset to the initial state (pressure 0,
    the initial saturation), so the timed steps are the workload's steps 1..K; each
    generic leg's saturation is compared with fused_fp64's after the same steps."""
    from equelle_amd.workloads import TWOPHASE, twophase_inputs, twophase_params
    nx, ny, nz = (int(v) for v in args.grid3.split("x"))
    n = nx * ny * nz
    perm, poro, sw0, sc, sv = twophase_inputs(nx, ny, nz, rate=TWOPHASE["rate"], porosity=TWOPHASE["porosity"])
    p0 = np.zeros(n)
    dt = TWOPHASE["dt"]
    runs = {}
    inst = f"watervisc={TWOPHASE['watervisc']} oilvisc={TWOPHASE['oilvisc']} corey={TWOPHASE['corey']}".encode()
    marks = sorted({g for _, _, g in generic_legs})   # fused_fp64 states kept after these step counts
    ref_at = {}

    def failed(name, err):
        log(f"[bench] two-phase {name}: {err}")
        runs[name] = {"failed": str(err)}

    legs = [(f"fused_{p}", p) for p in args.tp_legs.split(",")]
    if not legs or any(p not in ("fp64", "mixed") for _, p in legs) or legs[0][1] != "fp64":
        raise SystemExit("--tp-legs: fp64 first, then optionally mixed")
    for name, prec in legs:
        b = C.c_void_p()
        lib.batch_create(f"precision={prec} preconditioner={args.precond} rel_res_tol={TWOPHASE['rel_res_tol']} "
                         f"abs_res_tol=0 max_iter=20 solver_tol=1e-6 solver_max_iter=5000 "
                         f"graph={0 if args.no_graph else 1}".encode(), C.byref(b))
        try:
          try:
            slot = C.c_int()
            lib.batch_add_twophase(b, inst, nx, ny, nz, dptr(perm), dptr(poro), dptr(sw0), len(sc), iptr(sc),
                                   dptr(sv), C.byref(slot))
            sl = np.array([slot.value], dtype=np.int32)
            st, tps = np.zeros(5, dtype=np.int32), np.zeros(3, dtype=np.int32)
            for w_ in range(warmup):
                lib.batch_step(b, iptr(sl), 1, dt)
                lib.batch_stats(b, slot.value, iptr(st), None)
                lib.batch_twophase_stats(b, slot.value, iptr(tps))
                if st[3] == 0 or st[4] or tps[2] == 0:
                    raise RuntimeError(f"{name}: warmup step {w_} did not converge: {st.tolist()} {tps.tolist()}")
            lib.batch_set_twophase_state(b, slot.value, dptr(p0), dptr(sw0))
            lib.device_sync()
            newton = krylov = kp = 0
            excluded = 0.0
            lines = []   # printed after the timed steps
            t0 = time.perf_counter()
            for s_ in range(steps):
                lib.batch_step(b, iptr(sl), 1, dt)
                lib.batch_stats(b, slot.value, iptr(st), None)
                lib.batch_twophase_stats(b, slot.value, iptr(tps))
                if st[3] == 0 or st[4] or tps[2] == 0:
                    raise RuntimeError(f"{name}: step {s_} did not converge: {st.tolist()} {tps.tolist()}")
                newton += int(st[0])
                krylov += int(st[1])
                kp += int(tps[1])
                lines.append(f"[bench] two-phase {name} step {s_}: Newton p/s {tps[0]}/{st[0] - tps[0]}, "
                             f"Krylov p/s {tps[1]}/{st[1] - tps[1]}")
                if prec == "fp64" and s_ + 1 in marks:   # the states the generic legs end in
                    tr = time.perf_counter()
                    sw_m = np.zeros(n)
                    lib.batch_twophase_state(b, slot.value, None, dptr(sw_m))
                    ref_at[s_ + 1] = sw_m
                    excluded += time.perf_counter() - tr   # the copy is not part of the step
            lib.device_sync()
            secs = time.perf_counter() - t0 - excluded
            for ln in lines:
                log(ln)
            sw = np.zeros(n)
            lib.batch_twophase_state(b, slot.value, None, dptr(sw))
          except RuntimeError as e:   # FP32 Krylov on this pressure system: see the docstring
            if prec == "fp64":
                raise
            failed(name, e)
            continue
        finally:
            lib.batch_destroy(b)
        runs[name] = dict(value=n * newton / secs, ms_per_step=1e3 * secs / steps, steps=steps, newton=newton,
                          krylov_pressure=kp, krylov_transport=krylov - kp, sw=sw)
    for name, extra, gsteps in generic_legs:
        params = twophase_params(nx, ny, nz, rel_res_tol=TWOPHASE["rel_res_tol"], extra=extra).encode()
        sim = C.c_void_p()
        lib.twophase_create(params, dptr(perm), dptr(poro), dptr(sw0), n, iptr(sc), dptr(sv), len(sc), C.byref(sim))
        try:
          try:
            st = np.zeros(6, dtype=np.int32)
            for _ in range(min(warmup, 1)):   # one warmup step builds the symbolic cache
                lib.twophase_step(sim, dt, iptr(st), None)
            lib.twophase_set_state(sim, dptr(p0), dptr(sw0))
            lib.device_sync()
            newton, krylov = 0, [0, 0]
            t0 = time.perf_counter()
            for s_ in range(gsteps):
                lib.twophase_step(sim, dt, iptr(st), None)
                if st[2] == 0 or st[5] == 0:
                    raise RuntimeError(f"{name}: step {s_} did not converge: {st.tolist()}")
                newton += int(st[0]) + int(st[3])
                krylov[0] += int(st[1])
                krylov[1] += int(st[4])
                log(f"[bench] two-phase {name} step {s_}: Newton p/s {st[0]}/{st[3]}, Krylov p/s {st[1]}/{st[4]}")
            lib.device_sync()
            secs = time.perf_counter() - t0
            p, sw = np.zeros(n), np.zeros(n)
            lib.twophase_state(sim, dptr(p), dptr(sw))
          except RuntimeError as e:
            failed(name, e)
            continue
        finally:
            lib.twophase_destroy(sim)
        runs[name] = dict(value=n * newton / secs, ms_per_step=1e3 * secs / gsteps, newton=newton, steps=gsteps,
                          method=extra, krylov_pressure=krylov[0], krylov_transport=krylov[1], sw=sw)
    ref = runs["fused_fp64"]["sw"]
    errs = {}
    for k, r in runs.items():
        if "sw" not in r:
            continue
        want = ref_at[r["steps"]] if k.startswith("generic") else ref
        errs[k] = float(np.linalg.norm(r["sw"] - want) / np.linalg.norm(want))
    for r in runs.values():
        r.pop("sw", None)
    head = runs["fused_fp64"]
    out = {"value": head["value"], "unit": "cell-Newton-iters/s", "ms_per_step": head["ms_per_step"],
           "steps": steps, "cells": n, "runs": runs, "sw_rel_error_vs_fused_fp64": errs}
    if "value" in runs.get("fused_mixed", {}):
        out["fused_mixed_over_fused_fp64"] = runs["fused_mixed"]["value"] / head["value"]
    for name, _, _ in generic_legs:
        if "value" in runs.get(name, {}):
            out[f"fused_fp64_over_{name}"] = head["value"] / runs[name]["value"]
    return out


def bench_twophase(lib, args):
    """The two-phase workload on its own: one JSON line (value = fused FP64)."""
    from equelle_amd.workloads import TWOPHASE
    legs = [] if args.quick else [
        ("generic_mixed", f"precision=mixed solver={args.tp_solver} preconditioner=diagonal", args.steps),
        ("generic_fp64", f"precision=fp64 solver={args.tp_solver} preconditioner=diagonal", max(1, min(args.steps, 2)))]
    r = twophase_runs(lib, args, args.steps, args.warmup, legs)
    out = {"metric": "cell-Newton-iters/s", "value": r["value"], "unit": "cell-Newton-iters/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": r["ms_per_step"],
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"3D two-phase transport {args.grid3}, log-normal permeability, Corey 2, "
                                  f"viscosity ratio 5, dt={TWOPHASE['dt']}, Newton to 1e-8 relative residual per "
                                  "solve", "cells": r["cells"], "precond": args.precond}}
    out.update({k: v for k, v in r.items() if k not in ("value", "unit", "ms_per_step", "steps", "cells")})
    print(json.dumps(out), flush=True)


def bench_twophase_fi(lib, args):
    """Fully implicit two-phase flow (examples/dsl/twophase_fully_implicit.equelle): per
    step one newtonSolveSystem on (p, sw), i.e. one 2n x 2n Jacobian with the coupling
    blocks, on the device AD engine (the drop-in path; the fused engine has no fully
    implicit kind).  The workload's permeability, porosity and sources on a --grid3-fi
    grid, the program's linear relative permeabilities, dt = 0.05 (plain Newton from the
    program's initial guess sw = 0.5 converges there), each solve to 1e-8 relative
    residual; FP64 GMRES(30) with linear tolerance 1e-7 (the system is singular in the
    pressure constants; GMRES stays above its rounding floor), preconditioned by CPR (one
    FP32 multigrid V-cycle on the pressure block, then ILU(0) of the coupled Jacobian;
    --fi-precond ILU0 for ILU(0) alone).  Work = cells x
    Newton iterations.  cpu_baseline: the oracle (one core) on the same case, direct or
    BiCGStab/ILU(0) solves, one step."""
    from equelle_amd.workloads import twophase_inputs, twophase_params
    nx, ny, nz = (int(v) for v in args.grid3_fi.split("x"))
    n = nx * ny * nz
    perm, poro, sw0, sc, sv = twophase_inputs(nx, ny, nz)
    dt = 0.05
    params = twophase_params(nx, ny, nz, rel_res_tol=1e-8, max_iter=40, solver_tol=1e-7,
                             extra=f"solver=GMRes preconditioner={args.fi_precond} fully_implicit=1").replace(
                                 "corey=2", "corey=1")
    params = params.replace("solver=GMRes preconditioner=diagonal", "")
    sim = C.c_void_p()
    lib.twophase_create(params.encode(), dptr(perm), dptr(poro), dptr(sw0), n, iptr(sc), dptr(sv), len(sc),
                        C.byref(sim))
    st = np.zeros(6, dtype=np.int32)
    try:
        for _ in range(args.warmup):
            lib.twophase_step(sim, dt, iptr(st), None)
        lib.device_sync()
        newton = krylov = 0
        t0 = time.perf_counter()
        for s_ in range(args.steps):
            lib.twophase_step(sim, dt, iptr(st), None)
            if st[2] == 0:
                raise RuntimeError(f"fully implicit step {s_} did not converge: {st.tolist()}")
            newton += int(st[0])
            krylov += int(st[1])
            log(f"[bench] fully implicit step {s_}: Newton {st[0]}, GMRES {st[1]}")
        lib.device_sync()
        secs = time.perf_counter() - t0
    finally:
        lib.twophase_destroy(sim)
    value = n * newton / secs
    out = {"metric": "cell-Newton-iters/s", "value": value, "unit": "cell-Newton-iters/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * secs / args.steps,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"fully implicit two-phase {nx}x{ny}x{nz} (newtonSolveSystem, 2n unknowns), "
                                  "log-normal permeability, linear relative permeabilities, dt=0.05, Newton to 1e-8 "
                                  "relative residual; device AD engine, FP64 GMRES(30) with preconditioner "
                                  f"{args.fi_precond} (cpr: one FP32 multigrid V-cycle on the pressure block, then "
                                  "ILU(0) of the coupled system)",
                      "cells": n, "unknowns": 2 * n},
           "newton_iterations": newton, "krylov_iterations": krylov}
    if not args.no_cpu_baseline:
        oracle = CApi(os.path.join(ROOT, "oracle", "liboracle.so"), "eqo_")
        oparams = twophase_params(nx, ny, nz, rel_res_tol=1e-8, max_iter=40).replace("corey=2", "corey=1")
        p_o, s_o = np.zeros(n), np.zeros(n)
        it = np.zeros(1, dtype=np.int32)
        t = time.perf_counter()
        oracle.twophase_fi_run(oparams.encode(), dptr(perm), dptr(poro), dptr(sw0), n, iptr(sc), dptr(sv), len(sc),
                               dptr(np.array([dt])), 1, dptr(p_o), dptr(s_o), iptr(it))
        csecs = time.perf_counter() - t
        out["cpu_baseline"] = {"value": n * int(it[0]) / csecs, "unit": "cell-Newton-iters/s", "cores": 1,
                               "kind": "port",
                               "sample": f"first step of the same case ({int(it[0])} Newton iterations), serial C++ "
                                         f"oracle: {csecs:.2f} s"}
    print(json.dumps(out), flush=True)


def hip_runtime():
    """The HIP runtime library (libamdhip64), through ctypes."""
    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            return C.CDLL(name)
        except OSError:
            continue
    raise RuntimeError("libamdhip64.so not found")


def bench_copy(args):
    """Calibration for HBM counters: a device-to-device hipMemcpy of a known size (the HIP
    runtime's copy kernel), 256 MiB read + 256 MiB written per launch."""
    hip = hip_runtime()

    def ok(rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed with HIP error {rc}")

    a, b = C.c_void_p(), C.c_void_p()
    ok(hip.hipMalloc(C.byref(a), C.c_size_t(COPY_BYTES)), "hipMalloc")
    ok(hip.hipMalloc(C.byref(b), C.c_size_t(COPY_BYTES)), "hipMalloc")
    try:
        ok(hip.hipMemset(a, C.c_int(1), C.c_size_t(COPY_BYTES)), "hipMemset")
        for _ in range(args.warmup + args.steps):
            ok(hip.hipMemcpy(b, a, C.c_size_t(COPY_BYTES), C.c_int(3)), "hipMemcpy")   # device to device
        ok(hip.hipDeviceSynchronize(), "hipDeviceSynchronize")
    finally:
        hip.hipFree(a)
        hip.hipFree(b)
    print(json.dumps({"workload": "copy", "bytes_read_per_launch": COPY_BYTES,
                      "bytes_written_per_launch": COPY_BYTES, "launches": args.warmup + args.steps}), flush=True)


def pmc_per_dispatch(csv_path):
    """Kernel name -> counter values per dispatch (in dispatch order) from a rocprofv3
    counter-collection CSV of one counter."""
    import csv
    rows = list(csv.DictReader(open(csv_path)))
    if rows and "Dispatch_Id" in rows[0]:
        rows.sort(key=lambda r: int(r["Dispatch_Id"]))
    per = {}
    for r in rows:
        per.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]))
    return per


def measure_traffic(args):
    """HBM bytes per launch of every kernel of the headline loop, measured in this run:
    rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE (one counter per run, no other tracing)
    over child runs of this script (the headline at one timed step, then the 256 MiB copy),
    each counter scaled to bytes by what it reads for the copy.  Per kernel: the mean over
    the last 50 dispatches of its most-dispatched launch shape (the back-to-back timing
    launches on the final state).  The launch shape comes from the tuner of an unprofiled
    child run first; the counter runs and the headline run in it.  Runs before this process
    touches the GPU."""
    import glob
    import shutil
    import signal
    import subprocess
    import tempfile
    if shutil.which("rocprofv3") is None:
        return {"error": "rocprofv3 not on PATH"}
    me = os.path.abspath(__file__)
    head = ["--steps", "1", "--warmup", "0", "--no-cpu-baseline", "--quick", "--size", str(args.size),
            "--precond", args.precond, "--inner-tol", repr(args.inner_tol)]
    if args.forcing is not None:
        head += ["--forcing", repr(args.forcing)]
    if args.forcing0 is not None:
        head += ["--forcing0", repr(args.forcing0)]
    got = {}
    env = dict(os.environ, TMPDIR="/tmp")
    # the launch shape: the tuner's choice in an unprofiled child run (counter collection
    # distorts the tuner's timings); the counter runs and the headline use it
    shape = args.shape
    if shape is None:
        log("[bench] launch shape: tuner in an unprofiled child run")
        p = subprocess.Popen([sys.executable, me] + head, env=env, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                             start_new_session=True)
        try:
            sout, _ = p.communicate(timeout=240)
            line = json.loads([ln for ln in sout.decode(errors="replace").splitlines() if ln.startswith("{")][-1])
            cp, ls = line["launch_shape"]["cells_per_thread"], line["launch_shape"]
            shape = ",".join(str(cp[k]) for k in ("assemble", "defect", "spmv", "precond", "vector")) + \
                f",{ls['xcd_block_order']}"
        except subprocess.TimeoutExpired:
            os.killpg(p.pid, signal.SIGKILL)
            p.communicate()
        except (IndexError, KeyError, ValueError):
            pass
    with tempfile.TemporaryDirectory(dir="/tmp") as d:
        for leg, ctr in (("head", "FETCH_SIZE"), ("head", "WRITE_SIZE"), ("copy", "FETCH_SIZE"),
                         ("copy", "WRITE_SIZE")):
            extra = (head + (["--shape", shape] if shape else [])) if leg == "head" else \
                ["--workload", "copy", "--steps", "3", "--warmup", "1"]
            out = os.path.join(d, f"{leg}_{ctr}")
            cmd = ["rocprofv3", "--pmc", ctr, "--output-format", "csv", "-d", out, "-o", "run", "--",
                   sys.executable, me] + extra
            log(f"[bench] traffic: {ctr} over the {leg} run")
            p = subprocess.Popen(cmd, cwd=d, env=env, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                                 start_new_session=True)
            try:
                sout, err = p.communicate(timeout=240)
            except subprocess.TimeoutExpired:
                os.killpg(p.pid, signal.SIGKILL)
                p.communicate()
                return {"error": f"rocprofv3 --pmc {ctr} over the {leg} run timed out"}
            if leg == "head" and shape is None:
                try:
                    line = json.loads([ln for ln in sout.decode(errors="replace").splitlines()
                                       if ln.startswith("{")][-1])
                    cp, ls = line["launch_shape"]["cells_per_thread"], line["launch_shape"]
                    shape = ",".join(str(cp[k]) for k in ("assemble", "defect", "spmv", "precond", "vector")) + \
                        f",{ls['xcd_block_order']}"
                except (IndexError, KeyError, ValueError):
                    return {"error": "the first counter run printed no launch shape"}
            files = glob.glob(os.path.join(out, "**", "*counter_collection.csv"), recursive=True)
            if p.returncode != 0 and not files:
                return {"error": f"rocprofv3 --pmc {ctr} over the {leg} run exited {p.returncode}: "
                                 f"{err.decode(errors='replace')[-300:]}"}
            if not files:
                return {"error": f"rocprofv3 --pmc {ctr} over the {leg} run wrote no counter file"}
            got[(leg, ctr)] = pmc_per_dispatch(files[0])
    copy = {ctr: [v for k, v in got[("copy", ctr)].items() if "copy" in k.lower()] for ctr in ("FETCH_SIZE",
                                                                                               "WRITE_SIZE")}
    if not copy["FETCH_SIZE"] or not copy["WRITE_SIZE"]:
        return {"error": "no copy kernel in the calibration run"}
    kib_r = max(v[-1] for v in copy["FETCH_SIZE"])
    kib_w = max(v[-1] for v in copy["WRITE_SIZE"])
    scale_r, scale_w = COPY_BYTES / (kib_r * 1024), COPY_BYTES / (kib_w * 1024)
    per_kernel = {}
    fetch, write = got[("head", "FETCH_SIZE")], got[("head", "WRITE_SIZE")]
    for name, vals in fetch.items():
        if "eqhip" not in name or name not in write:
            continue
        short = name.split("(eqhip")[0].replace("void ", "").replace("eqhip::(anonymous namespace)::", "")
        f, w = vals[-50:], write[name][-50:]
        per_kernel[short] = {"dispatches": len(vals), "read_bytes": sum(f) / len(f) * 1024 * scale_r,
                             "write_bytes": sum(w) / len(w) * 1024 * scale_w}
    return {"per_kernel": per_kernel, "scale_read": scale_r, "scale_write": scale_w, "shape": shape,
            "source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE child runs of this script (headline at 1 timed "
                      "step; 256 MiB device-to-device copy for the byte scale); the counter runs and the "
                      "headline of this run use the launch shape an unprofiled child run tuned"}


def traffic_of(measured, kernel):
    """Measured bytes per launch of `kernel` (e.g. "k_schur_black<float, 4>"): the
    most-dispatched launch shape with the same name and first template argument."""
    if not measured or "per_kernel" not in measured:
        return None
    base, _, rest = kernel.partition("<")
    first = rest.split(",")[0].rstrip(">") if "," in rest else None
    cands = [(v["dispatches"], k, v) for k, v in measured["per_kernel"].items()
             if k.startswith(base + "<") and (first is None or k[len(base) + 1:].split(",")[0] == first)]
    if not cands:
        return None
    _, k, v = max(cands)
    return {"kernel": k, "bytes": v["read_bytes"] + v["write_bytes"], "read_bytes": v["read_bytes"],
            "write_bytes": v["write_bytes"]}


# Algorithmic bytes per launch over n cells (h = (n + 1) / 2 pair ranks; ks = bytes of the
# Krylov precision): every array the kernel must touch, once (stencil neighbours are
# re-reads of the same array).  Launches per timed region from the iteration counts:
# every BiCGStab iteration applies the reduced operator twice (both passes) and the three
# vector updates once; every refinement round loads a right-hand side and runs a defect;
# every Newton iteration and every step start assembles.
KERNELS = {
    "operator": ("k_schur_vec<T, {cpt}>", "spmv", lambda n, h, ks: 13 * ks * h,
                 "one pass (y_r formed in registers): red 4 coefficients, 1/C_r; black C_b, 4 coefficients; "
                 "v_b, dot operand, w written (13 ks h)"),
    "operator_red": ("k_schur_red<T, {cpt}>", "spmv", lambda n, h, ks: 7 * ks * h,
                     "red ranks: 4 coefficients, 1/C_r, v_b, y_r written (7 ks h)"),
    "operator_black": ("k_schur_black<T, {cpt}>", "spmv", lambda n, h, ks: 9 * ks * h,
                       "black ranks: C_b, 4 coefficients, v_b, y_r, dot operand, w written (9 ks h)"),
    "spmv": ("k_spmv_rb<T, {cpt}>", "spmv", lambda n, h, ks: 2 * 9 * ks * h,
             "both colours: C, 4 coefficients, x, dot operand, y written (18 ks h)"),
    "assemble": ("k_assemble<{mixed}, {cpt}>", "assemble", lambda n, h, ks: 24 * n + 5 * ks * 2 * h,
                 "u, u0 read, F written (FP64), 5 coefficients written (Krylov precision)"),
    "defect": ("k_defect<{cpt}>", "defect", lambda n, h, ks: 32 * n, "u, du, F read, r written (FP64)"),
    "b1": ("k_b1<T, {cpt}>", "vector", lambda n, h, ks: 6 * ks * h, "r, p, v, 1/P read; p, ph written"),
    "b2": ("k_b2<T, {cpt}>", "vector", lambda n, h, ks: 5 * ks * h, "r, v, 1/P read; s, sh written"),
    "b3": ("k_b3<T, {cpt}>", "vector", lambda n, h, ks: 8 * ks * h, "x, ph, sh, s, t, rh read; x, r written"),
    "b31": ("k_b31<T, {cpt}>", "vector", lambda n, h, ks: 12 * ks * h,
            "x, ph, sh, s, t, p, v, 1/P read; x, r, p, ph written"),
    "load_rhs": ("k_load_rhs<T, {cpt}>", "precond", lambda n, h, ks: 8 * n + 6 * ks * h,
                 "r64 read (FP64), 4 black coefficients, 1/C_r read, g written"),
    "setup": ("k_setup_vec<T, {cpt}>", "precond", lambda n, h, ks: 8 * n + 17 * ks * h,
              "r64 read (FP64), 10 coefficient arrays read; 1/C_r, 1/P_b, r, rh, p, ph, x written"),
}


# The four-ranks-per-lane form solves the column-scaled reduced system (the preconditioner
# lives in the black columns): its vector kernels carry no preconditioned vectors.
KERNELS_VEC = {
    "b2": (lambda n, h, ks: 3 * ks * h, "r, v read; s written"),
    "b31": (lambda n, h, ks: 8 * ks * h, "y, p, s, t, v read; y, r, p written"),
    # the timed launches leave the coefficients alone (no pivots, no column scaling: those
    # writes, 1/P_b and the 5 scaled black columns, 6 ks h more, happen in a Newton
    # iteration's first refinement round only)
    "setup": (lambda n, h, ks: 8 * n + 15 * ks * h,
              "r64 (FP64) and the 10 coefficient arrays read; 1/C_r, r, rh, p, y written (timed without the "
              "first round's pivot and column-scaling writes)"),
}


def kernel_table(run, n, ks, steps):
    h = (n + 1) // 2
    kern = run["kern"]
    per = {"operator": 2 * run["krylov"], "operator_red": 2 * run["krylov"], "operator_black": 2 * run["krylov"],
           "spmv": 2 * run["krylov"],
           "b1": run["krylov"], "b2": run["krylov"], "b3": run["krylov"], "b31": run["krylov"],
           "assemble": run["newton"] + steps,
           "defect": run["refine"], "load_rhs": run["refine"] if ks == 4 else run["newton"],
           "setup": run["refine"] if ks == 4 else run["newton"]}
    out = {}
    for name, (kname, cls, fb, text) in KERNELS.items():
        if name not in kern:
            continue
        if name == "operator" and kern.get("tiles"):   # nx % 512 == 0: the tiled form of the operator
            kname = "k_schur_tile<T>"
        if kern.get("vec") and name in KERNELS_VEC:   # the column-scaled four-ranks-per-lane form
            fb, text = KERNELS_VEC[name]
        _, ms = kern[name]
        bytes_ = fb(n, h, ks)
        gbs = bytes_ / (ms * 1e-3) / 1e9
        out[name] = {"kernel": kname.format(cpt=kern["cpt"][cls], mixed=str(ks == 4).lower()).replace(
                         "T", "float" if ks == 4 else "double"),
                     "avg_us": 1e3 * ms, "back_to_back_us": 1e3 * kern[name + "_event"],
                     "algorithmic_bytes": bytes_, "bytes": text, "GBs": gbs, "frac": gbs / HBM_PEAK_GBS,
                     "us_per_step": 1e3 * ms * per[name] / steps}
    return out


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
                cpu_model=cpu_model(),
                sample=f"1 Newton iteration (AD residual+Jacobian, BiCGStab/ILU(0) to 1e-6, {lin.value} "
                       f"iterations) of the {n_side}x{n_side} nonlinear-heat step, serial C++ oracle: "
                       f"{secs.value:.2f} s")


def cpu_model():
    """The host CPU's model name and logical CPU count (the box's own /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            names = [ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")]
        return f"{names[0]} ({len(names)} logical CPUs visible)" if names else "unknown"
    except OSError:
        return "unknown"


def serving_compact(lib, args):
    """The request loop at two arrival rates: latency percentiles per request kind."""
    from equelle_amd.serving import serving_sweep
    sweep = serving_sweep(lib, [float(r) for r in args.default_rates.split(",")], args.default_requests,
                          frac3d=args.frac3d, precond=args.precond, hbm_budget_gb=args.hbm_budget_gb)
    keep = ("rate_req_s", "requests", "p50_ms", "p99_ms", "heat2d", "twophase3d", "throughput_req_s",
            "cell_newton_iters_per_s", "failed")
    return {"workload": f"Poisson stream of {args.default_requests} requests per rate, {100 * (1 - args.frac3d):.0f}% "
                        f"2D heat (16..1024), {100 * args.frac3d:.0f}% 3D two-phase (16..128), 1-20 steps each",
            "sweep": [{k: s[k] for k in keep if k in s} for s in sweep]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size", type=int, default=NONLINEAR_HEAT["nx"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--quick", action="store_true", help="headline run only (for profiling)")
    ap.add_argument("--no-traffic", dest="no_traffic", action="store_true",
                    help="skip the rocprofv3 counter runs that measure roofline.traffic")
    ap.add_argument("--shape", default=None,
                    help="headline launch shape a,d,s,p,v,xcd (cells per thread of assembly, defect, SpMV, "
                         "preconditioner and vector kernels; XCD block order) instead of the tuner's")
    ap.add_argument("--no-extras", dest="no_extras", action="store_true",
                    help="heat headline legs only (no batched, two-phase and serving summaries)")
    ap.add_argument("--default-rates", dest="default_rates", default="20,50,500",
                    help="serving rates of the default run's compact summary")
    ap.add_argument("--default-requests", dest="default_requests", type=int, default=100)
    ap.add_argument("--precond", default="ilu0", help="preconditioner of the fused runs: ilu0 | diagonal | none")
    ap.add_argument("--forcing", type=float, default=None,
                    help="fused heat runs: inexact-Newton margin of the linear tolerance (0: solver_tol always)")
    ap.add_argument("--forcing0", type=float, default=None,
                    help="fused heat runs: first-iteration factor on the latest step's contraction (0: solver_tol)")
    ap.add_argument("--workload", default="heat2048",
                    choices=["heat2048", "batched", "serving", "twophase", "twophase_fi", "copy", "generic"])
    ap.add_argument("--instances", type=int, default=512)
    ap.add_argument("--inner-tol", dest="inner_tol", type=float, default=1e-6,
                    help="mixed precision: FP32 inner solve tolerance")
    ap.add_argument("--no-graph", dest="no_graph", action="store_true", help="eager launches (profiling)")
    ap.add_argument("--grid3", default="128x128x64", help="two-phase grid nx x ny x nz")
    ap.add_argument("--grid3-fi", dest="grid3_fi", default="64x64x32", help="fully implicit two-phase grid")
    ap.add_argument("--fi-precond", dest="fi_precond", default="cpr",
                    help="fully implicit: GMRES preconditioner (ILU0 | cpr | diagonal)")
    ap.add_argument("--tp-solver", dest="tp_solver", default="GMRes", help="two-phase Krylov solver")
    ap.add_argument("--tp-legs", dest="tp_legs", default="fp64,mixed",
                    help="two-phase fused legs: fp64 (the headline) and optionally mixed")
    ap.add_argument("--rates", default="50,100,200,500,1000,2000", help="serving: arrival rates (req/s)")
    ap.add_argument("--requests", type=int, default=200, help="serving: requests per rate")
    ap.add_argument("--frac3d", type=float, default=0.3, help="serving: share of 3D two-phase requests")
    ap.add_argument("--hbm-budget-gb", dest="hbm_budget_gb", type=float, default=64.0)
    ap.add_argument("--max-in-flight-twophase", dest="max_in_flight_twophase", type=int, default=4096,
                    help="serving: two-phase requests in flight at once")
    args = ap.parse_args()
    global FORCING
    if args.forcing is not None:
        FORCING = f"forcing={args.forcing}"
    if args.forcing0 is not None:
        FORCING += f" forcing0={args.forcing0}"
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "copy":
        return bench_copy(args)
    # the counter runs are child processes started before this process touches the GPU
    measured = measure_traffic(args) if args.workload == "heat2048" and not (args.quick or args.no_traffic) \
        else None
    lib = load_library()
    if args.workload == "batched":
        return bench_batched(lib, args)
    if args.workload == "twophase":
        return bench_twophase(lib, args)
    if args.workload == "twophase_fi":
        return bench_twophase_fi(lib, args)
    if args.workload == "generic":
        return bench_generic(lib, args)
    if args.workload == "serving":
        from equelle_amd.serving import bench_serving
        return bench_serving(lib, args)
    nx = ny = args.size
    n = nx * ny
    nif = (nx - 1) * ny + nx * (ny - 1)
    nnz = n + 2 * nif
    dt = NONLINEAR_HEAT["dt"]
    u0 = gaussian_bumps(nx, ny, seed=2048)
    log(f"[bench] nonlinear heat {nx}x{ny}: fused engine, mixed precision")
    shape = args.shape or (measured or {}).get("shape")
    head = run_fused(lib, "mixed", u0, nx, ny, dt, args.steps, args.warmup, args.precond, not args.no_graph,
                     args.inner_tol, extra=f"shape={shape}" if shape else "")
    value = n * head["newton"] / head["seconds"]
    kt = kernel_table(head, n, 4, args.steps)
    dom = max(kt, key=lambda k: kt[k]["us_per_step"])
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
        "config": {"workload": f"nonlinear heat {nx}x{ny}, k(u)=0.1(1+u^2), dt=1, zero-flux, Newton to 1e-8 "
                               "relative residual; mixed precision (FP32 BiCGStab on the red-black reduced "
                               "system with red-black ILU(0) pivots, inner tolerance "
                               f"{args.inner_tol:g}, FP64 refinement to 1e-6)",
                   "cells": n, "jacobian_nnz": nnz, "engine": "fused heat engine (1 instance)"},
        "roofline": {"bound": "hbm", "achieved": kt[dom]["GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": kt[dom]["frac"],
                     "traffic": None,
                     "kernel": kt[dom]["kernel"], "launches": 50, "avg_us": kt[dom]["avg_us"],
                     "back_to_back_us": kt[dom]["back_to_back_us"],
                     "algorithmic_bytes": kt[dom]["algorithmic_bytes"], "bytes_formula": kt[dom]["bytes"]},
        "assembly_roofline": kt["assemble"],
        "kernels": kt,
        "launch_shape": {"cells_per_thread": head["kern"]["cpt"], "xcd_block_order": head["kern"]["shape"][1],
                         "chosen_by": "launch-shape tuner (each kernel class timed per batch membership)"},
        "newton_iterations": head["newton"],
        "krylov_iterations": head["krylov"],
        "cell_steps_per_s": n * args.steps / head["seconds"],
    }
    tr = traffic_of(measured, kt[dom]["kernel"])
    if tr is not None:
        out["roofline"]["traffic"] = tr["bytes"]
        out["roofline"]["traffic_detail"] = dict(tr, scale_read=measured["scale_read"],
                                                 scale_write=measured["scale_write"], source=measured["source"],
                                                 over_algorithmic=tr["bytes"] / kt[dom]["algorithmic_bytes"])
        for name, row in kt.items():
            t = traffic_of(measured, row["kernel"])
            if t is not None:
                row["traffic"] = t["bytes"]
    elif measured is not None:
        out["roofline"]["traffic_detail"] = {"error": measured.get("error", "kernel not in the counter runs")}
    if not args.quick:
        log("[bench] plain FP64 (fused engine)")
        p = run_fused(lib, "fp64", u0, nx, ny, dt, args.steps, args.warmup, args.precond, not args.no_graph)
        pv = n * p["newton"] / p["seconds"]
        kt64 = kernel_table(p, n, 8, args.steps)
        out["plain_fp64"] = {"value": pv, "ms_per_step": 1e3 * p["seconds"] / args.steps,
                             "newton_iterations": p["newton"], "krylov_iterations": p["krylov"],
                             "kernels": {k: {"avg_us": v["avg_us"], "frac": v["frac"]} for k, v in kt64.items()}}
        out["mixed_over_plain"] = value / pv
        out["mixed_vs_fp64_rel_error"] = float(np.linalg.norm(head["u"] - p["u"]) / np.linalg.norm(p["u"]))
        log("[bench] generic drop-in path (EquelleRuntimeCUDA AD engine, FP64)")
        gsteps = max(2, args.steps // 4)
        g = run_generic(lib, u0, nx, ny, dt, gsteps, 1, eval_reps=10)
        out["generic_fp64"] = {"value": n * g["newton"] / g["seconds"], "steps": gsteps,
                               "ms_per_step": 1e3 * g["seconds"] / gsteps}
        out["generic_assembly_roofline"] = g["assembly"]
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline (serial oracle, one Newton iteration)")
        out["cpu_baseline"] = cpu_baseline(nx, u0, dt)
    if not args.quick and not args.no_extras:
        # the other configured workloads, compact: batched (512 instances), the 3D two-phase
        # workload's 30 steps with its configured method beside the fused engine, the request loop
        from equelle_amd.workloads import TWOPHASE
        w = min(args.warmup, 2)
        out["batched"] = batched_runs(lib, args, args.steps, w, 2)
        out["twophase"] = twophase_runs(
            lib, args, TWOPHASE["steps"], w,
            [("generic_configured", "precision=mixed solver=GMRes preconditioner=diagonal", TWOPHASE["steps"])])
        out["twophase"]["configured_method"] = ("generic_configured: FP64 residual/Jacobian, FP32 GMRES(30) + "
                                                "Jacobi, FP64 iterative refinement (the drop-in AD engine)")
        out["serving"] = serving_compact(lib, args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
