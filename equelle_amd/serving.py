"""Client side of the request loop (eqhip_server_*): a seeded Poisson request stream
replayed against the native server in wall-clock time, with latency percentiles.

Request mix (each request 1..20 time steps, every Newton solve to 1e-8 relative residual):
* 70% 2D nonlinear heat (the physics of the 2048^2 workload with k(u) = k0 (1 + u^2)):
  side ~ round(exp(U[ln 16, ln 1024])), k0 ~ U[0.05, 0.5], dt = 1, Gaussian-bump initial
  field scaled to the grid;
* 30% 3D two-phase transport (the two-phase workload's physics: log-normal permeability,
  Corey 2, viscosity ratio 5, injector and producer columns): side ~ round(exp(U[ln 16,
  ln 128])) cells in each direction, dt = 1.
"""
from __future__ import annotations

import ctypes as C
import json
import sys
import time
from dataclasses import dataclass

import numpy as np

from .capi import dptr, iptr
from .workloads import gaussian_bumps


@dataclass
class Request:
    t: float          # arrival time (s after the start of the stream)
    side: int
    k0: float
    steps: int
    u0: np.ndarray    # heat: initial field; two-phase: initial saturation
    dt: float = 1.0
    twophase: bool = False
    perm: np.ndarray = None
    poro: np.ndarray = None
    src_cells: np.ndarray = None
    src_vals: np.ndarray = None

    @property
    def n(self):
        return self.side ** 3 if self.twophase else self.side ** 2


TWOPHASE_INSTANCE = "watervisc=0.001 oilvisc=0.005 corey=2"


def twophase_request(t, side, steps, seed):
    from .workloads import twophase_inputs
    perm, poro, sw0, sc, sv = twophase_inputs(side, side, side, seed=seed)
    return Request(t, side, 0.0, steps, sw0, twophase=True, perm=perm, poro=poro, src_cells=sc, src_vals=sv)


def poisson_stream(rate: float, count: int, seed: int = 7, lo: int = 16, hi: int = 1024, max_steps: int = 20,
                   frac3d: float = 0.0, lo3: int = 16, hi3: int = 128):
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, count))
    sides = np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), count))).astype(int)
    sides3 = np.rint(np.exp(rng.uniform(np.log(lo3), np.log(hi3), count))).astype(int)
    is3 = rng.uniform(0.0, 1.0, count) < frac3d
    k0 = rng.uniform(0.05, 0.5, count)
    steps = rng.integers(1, max_steps + 1, count)
    out = []
    for i in range(count):
        if is3[i]:
            out.append(twophase_request(float(t[i]), int(sides3[i]), int(steps[i]), seed * 100003 + i))
        else:
            out.append(Request(float(t[i]), int(sides[i]), float(k0[i]), int(steps[i]),
                               gaussian_bumps(int(sides[i]), int(sides[i]), seed * 100003 + i,
                                              width_scale=sides[i] / 2048.0)))
    return out


class Server:
    """Owns one eqhip server handle."""

    def __init__(self, lib, params: str):
        self.lib = lib
        self.h = C.c_void_p()
        lib.server_create(params.encode(), C.byref(self.h))
        self._fin = np.zeros(4096, dtype=np.int32)

    def submit(self, r: Request) -> int:
        rid = C.c_int()
        if r.twophase:
            self.lib.server_submit_twophase(self.h, TWOPHASE_INSTANCE.encode(), r.side, r.side, r.side, dptr(r.perm),
                                            dptr(r.poro), dptr(r.u0), len(r.src_cells), iptr(r.src_cells),
                                            dptr(r.src_vals), r.steps, r.dt, C.byref(rid))
        else:
            self.lib.server_submit(self.h, r.side, r.side, r.k0, 1.0, dptr(r.u0), r.steps, r.dt, C.byref(rid))
        return rid.value

    def set_max_newton(self, rid: int, max_newton: int):
        self.lib.server_set_max_newton(self.h, rid, max_newton)

    def progress(self):
        nf = C.c_int()
        self.lib.server_progress(self.h, iptr(self._fin), len(self._fin), C.byref(nf))
        return self._fin[:nf.value].tolist()

    def result(self, rid: int, n: int, pressure=None):
        u = np.zeros(n)
        st = np.zeros(4, dtype=np.int32)
        tm = np.zeros(2)
        self.lib.server_result(self.h, rid, dptr(u), dptr(pressure) if pressure is not None else None, iptr(st),
                               dptr(tm))
        return u, st, tm

    def status(self):
        c = np.zeros(2, dtype=np.int32)
        b = np.zeros(2)
        self.lib.server_status(self.h, iptr(c), dptr(b))
        return c, b

    def close(self):
        if self.h:
            self.lib.server_destroy(self.h)
            self.h = C.c_void_p()


def replay(server: Server, stream, keep_results=False):
    """Submits each request at its arrival time and ticks the server until all finish.
    Latency = completion time - scheduled arrival time (includes client-side lag)."""
    out = {}
    pending = {}
    i = 0
    t0 = time.perf_counter()
    while i < len(stream) or pending:
        now = time.perf_counter() - t0
        while i < len(stream) and stream[i].t <= now:
            pending[server.submit(stream[i])] = i
            i += 1
        if pending:
            done = server.progress()
            tnow = time.perf_counter() - t0
            for rid in done:
                q = pending.pop(rid)
                r = stream[q]
                u, st, tm = server.result(rid, r.n)
                out[q] = dict(latency=tnow - r.t, server_latency=float(tm[1]), queued=float(tm[0]),
                              steps=int(st[0]), newton=int(st[1]), failed=bool(st[3]), twophase=r.twophase,
                              cell_newton=int(st[1]) * r.n, u=u if keep_results else None)
        elif i < len(stream):
            time.sleep(max(0.0, min(stream[i].t - now, 0.002)))
    return out, time.perf_counter() - t0


def summarize(stream, out, wall):
    lat = np.array([out[q]["latency"] for q in range(len(stream))])
    res = dict(requests=len(stream), wall_s=wall,
               throughput_req_s=len(stream) / wall,
               cell_newton_iters_per_s=sum(o["cell_newton"] for o in out.values()) / wall,
               p50_ms=1e3 * float(np.percentile(lat, 50)), p99_ms=1e3 * float(np.percentile(lat, 99)),
               mean_queue_ms=1e3 * float(np.mean([o["queued"] for o in out.values()])),
               failed=int(sum(o["failed"] for o in out.values())))
    for kind, sel in (("heat2d", False), ("twophase3d", True)):
        l = np.array([out[q]["latency"] for q in range(len(stream)) if stream[q].twophase == sel])
        if l.size:
            res[kind] = dict(requests=int(l.size), p50_ms=1e3 * float(np.percentile(l, 50)),
                             p99_ms=1e3 * float(np.percentile(l, 99)))
    return res


def serving_sweep(lib, rates, count, frac3d=0.3, precond="ilu0", hbm_budget_gb=64.0, cap=4096):
    """Replays a seeded Poisson stream of `count` requests at each rate; one summary per rate
    (latency percentiles overall and per request kind, throughput, failures)."""
    params = (f"precision=mixed preconditioner={precond} rel_res_tol=1e-8 abs_res_tol=0 max_iter=20 "
              f"solver_tol=1e-6 hbm_budget={hbm_budget_gb * 1e9} max_in_flight_twophase={cap}")
    sweep = []
    for rate in rates:
        stream = poisson_stream(rate, count, seed=7, frac3d=frac3d)
        srv = Server(lib, params)
        try:
            # warm the kernels (first launches) outside the measured stream
            srv.submit(Request(0.0, 64, 0.1, 1, gaussian_bumps(64, 64, 1, width_scale=64 / 2048)))
            while srv.progress() == [] or srv.status()[0][1]:
                pass
            out, wall = replay(srv, stream)
        finally:
            srv.close()
        s = summarize(stream, out, wall)
        s["rate_req_s"] = rate
        print(f"[serving] rate {rate:g}/s: p50 {s['p50_ms']:.1f} ms, p99 {s['p99_ms']:.1f} ms, "
              f"{s['throughput_req_s']:.1f} req/s, {s['cell_newton_iters_per_s']:.3g} cell-Newton/s",
              file=sys.stderr, flush=True)
        sweep.append(s)
    return sweep


def bench_serving(lib, args):
    """Rate sweep of the request loop; one JSON line (value = sustained cell-Newton-iters/s
    at the highest rate, latency percentiles per rate under `sweep`)."""
    rates = [float(r) for r in getattr(args, "rates", "50,100,200,500,1000,2000").split(",")]
    count = getattr(args, "requests", 200)
    frac3d = getattr(args, "frac3d", 0.3)
    cap = getattr(args, "max_in_flight_twophase", 4096)
    sweep = serving_sweep(lib, rates, count, frac3d, args.precond, getattr(args, "hbm_budget_gb", 64.0), cap)
    top = sweep[-1]
    line = {"metric": "cell-Newton-iters/s (serving)", "value": top["cell_newton_iters_per_s"],
            "unit": "cell-Newton-iters/s", "n_gpus": 1, "steps": count, "warmup": 1,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f64 residual/Jacobian; heat: f32 Krylov + f64 refinement, two-phase: f64 Krylov",
            "data": "synthetic",
            "config": {"workload": f"Poisson request stream: {100 * (1 - frac3d):.0f}% 2D nonlinear heat (side "
                                   f"log-uniform 16..1024), {100 * frac3d:.0f}% 3D two-phase (side log-uniform "
                                   "16..128), 1-20 steps, dt=1, Newton to 1e-8 relative residual",
                       "requests_per_rate": count, "hbm_budget_gb": getattr(args, "hbm_budget_gb", 64.0),
                       "max_in_flight_twophase": cap},
            "p50_ms": top["p50_ms"], "p99_ms": top["p99_ms"], "sweep": sweep}
    print(json.dumps(line), flush=True)
