"""The request loop (eqhip_server_*): admission by HBM footprint, continuous batching,
and the result of every request identical alone and in the mix.

The server's admitted footprint never exceeds the budget; requests that arrive while
others run join the batch between Newton iterations; each request's final field equals,
bit for bit, the same request run alone through the batch API (per-instance reductions
are deterministic and independent of the other slots), and matches the oracle's
newtonSolve trajectory to the solve tolerance.
"""
import ctypes as C

import numpy as np
import pytest

from equelle_amd.capi import EquelleError, dptr, iptr
from equelle_amd.serving import Request, Server, poisson_stream, replay
from equelle_amd.workloads import gaussian_bumps

pytestmark = pytest.mark.gpu

PARAMS = "precision=mixed preconditioner=ilu0 rel_res_tol=1e-8 abs_res_tol=0 max_iter=20 solver_tol=1e-6"


def alone(api, r):
    """The request run alone through the batch API (two-phase: the request loop's FP64
    Krylov precision, saturation returned)."""
    from equelle_amd.serving import TWOPHASE_INSTANCE
    b = C.c_void_p()
    params = PARAMS.replace("precision=mixed", "precision=fp64") if r.twophase else PARAMS
    api.batch_create(params.encode(), C.byref(b))
    try:
        s = C.c_int()
        if r.twophase:
            api.batch_add_twophase(b, TWOPHASE_INSTANCE.encode(), r.side, r.side, r.side, dptr(r.perm), dptr(r.poro),
                                   dptr(r.u0), len(r.src_cells), iptr(r.src_cells), dptr(r.src_vals), C.byref(s))
        else:
            api.batch_add(b, r.side, r.side, r.k0, 1.0, dptr(r.u0), C.byref(s))
        for _ in range(r.steps):
            api.batch_step(b, iptr(np.array([s.value], dtype=np.int32)), 1, r.dt)
        u = np.zeros(r.n)
        api.batch_state(b, s.value, dptr(u))
        return u
    finally:
        api.batch_destroy(b)


def test_mix_equals_alone_and_budget_respected(hip_api):
    reqs = [Request(0.0, side, k, steps, gaussian_bumps(side, side, 40 + q, width_scale=side / 512))
            for q, (side, k, steps) in enumerate([(96, 0.2, 3), (40, 0.05, 1), (150, 0.45, 2), (23, 0.3, 4),
                                                  (64, 0.1, 2), (128, 0.25, 1)])]
    # budget for about two of the larger requests at once: the rest must queue
    budget = 2.2 * 150 * 150 * 150
    srv = Server(hip_api, PARAMS + f" hbm_budget={budget}")
    try:
        ids = {}
        results = {}
        ticks = 0
        nxt = 0
        while len(results) < len(reqs):
            if nxt < len(reqs) and ticks % 2 == 0:       # arrivals interleaved with Newton iterations
                ids[srv.submit(reqs[nxt])] = nxt
                nxt += 1
            counts, used = srv.status()
            assert used[0] <= used[1]
            for rid in srv.progress():
                q = ids[rid]
                u, st, tm = srv.result(rid, reqs[q].n)
                assert st[0] == reqs[q].steps and st[3] == 0, st
                assert 0 <= tm[0] <= tm[1]
                results[q] = u
            ticks += 1
            assert ticks < 2000
        for q, r in enumerate(reqs):
            assert np.array_equal(results[q], alone(hip_api, r)), q
    finally:
        srv.close()


def test_serving_matches_oracle(hip_api, oracle_api):
    r = Request(0.0, 37, 0.3, 2, gaussian_bumps(37, 37, 5, width_scale=0.05))
    srv = Server(hip_api, PARAMS.replace("rel_res_tol=1e-8", "rel_res_tol=1e-12").replace("solver_tol=1e-6",
                                                                                           "solver_tol=1e-12"))
    try:
        rid = srv.submit(r)
        other = srv.submit(Request(0.0, 80, 0.1, 3, gaussian_bumps(80, 80, 6, width_scale=0.1)))
        got = None
        for _ in range(500):
            for f in srv.progress():
                u, st, _ = srv.result(f, 37 * 37 if f == rid else 80 * 80)
                if f == rid:
                    got = u
            if got is not None and srv.status()[0][1] == 0:
                break
        assert got is not None
    finally:
        srv.close()
    params = (b"grid_dim=2 nx=37 ny=37 k=0.3 cv=1 k_model=quadratic abs_res_tol=0 rel_res_tol=1e-12 max_iter=30")
    e, ed = np.zeros(1, dtype=np.int32), np.zeros(1)
    want = np.zeros(37 * 37)
    oracle_api.heat_run(params, dptr(r.u0), 37 * 37, iptr(e), dptr(ed), 0, dptr(np.array([1.0, 1.0])), 2, dptr(want),
                        None, None)
    assert np.allclose(got, want, rtol=1e-9, atol=1e-11)


def test_request_larger_than_budget_is_rejected(hip_api):
    srv = Server(hip_api, PARAMS + " hbm_budget=1e5")
    try:
        with pytest.raises(EquelleError, match="exceeds the HBM budget"):
            srv.submit(Request(0.0, 512, 0.1, 1, np.zeros(512 * 512)))
    finally:
        srv.close()


def test_poisson_replay_completes(hip_api):
    stream = poisson_stream(400.0, 24, seed=3, lo=16, hi=128, max_steps=4)
    srv = Server(hip_api, PARAMS)
    try:
        out, wall = replay(srv, stream)
    finally:
        srv.close()
    assert len(out) == len(stream) and not any(o["failed"] for o in out.values())
    assert all(o["steps"] == stream[q].steps for q, o in out.items())


def test_mixed_heat_and_twophase_requests(hip_api):
    """3D two-phase and 2D heat requests in one request loop (heat in mixed precision,
    two-phase with FP64 Krylov): each result equals the same request run alone through
    the batch API with that precision, bit for bit."""
    from equelle_amd.serving import TWOPHASE_INSTANCE, twophase_request
    reqs = [Request(0.0, 50, 0.2, 2, gaussian_bumps(50, 50, 60, width_scale=0.03)),
            twophase_request(0.0, 12, 2, seed=61),
            Request(0.0, 90, 0.4, 3, gaussian_bumps(90, 90, 62, width_scale=0.05)),
            twophase_request(0.0, 9, 3, seed=63)]
    srv = Server(hip_api, PARAMS)
    try:
        ids = {}
        results = {}
        for q, r in enumerate(reqs):
            ids[srv.submit(r)] = q
            srv.progress()   # later requests join mid-flight
        for _ in range(2000):
            for rid in srv.progress():
                q = ids[rid]
                p = np.zeros(reqs[q].n)
                u, st, _ = srv.result(rid, reqs[q].n, pressure=p)
                assert st[0] == reqs[q].steps and st[3] == 0, st
                results[q] = (u, p)
            if len(results) == len(reqs):
                break
        assert len(results) == len(reqs)
    finally:
        srv.close()
    for q, r in enumerate(reqs):
        b = C.c_void_p()
        # the request loop runs two-phase requests with FP64 Krylov (twophase_precision=fp64)
        params = PARAMS.replace("precision=mixed", "precision=fp64") if r.twophase else PARAMS
        hip_api.batch_create(params.encode(), C.byref(b))
        try:
            s = C.c_int()
            if r.twophase:
                hip_api.batch_add_twophase(b, TWOPHASE_INSTANCE.encode(), r.side, r.side, r.side, dptr(r.perm),
                                           dptr(r.poro), dptr(r.u0), len(r.src_cells), iptr(r.src_cells),
                                           dptr(r.src_vals), C.byref(s))
            else:
                hip_api.batch_add(b, r.side, r.side, r.k0, 1.0, dptr(r.u0), C.byref(s))
            for _ in range(r.steps):
                hip_api.batch_step(b, iptr(np.array([s.value], dtype=np.int32)), 1, r.dt)
            u = np.zeros(r.n)
            hip_api.batch_state(b, s.value, dptr(u))
            assert np.array_equal(results[q][0], u), q
            if r.twophase:
                p = np.zeros(r.n)
                hip_api.batch_twophase_state(b, s.value, dptr(p), None)
                assert np.array_equal(results[q][1], p), q
        finally:
            hip_api.batch_destroy(b)


def test_failed_requests_are_isolated(hip_api):
    """A request with a NaN in its initial field and a request limited to one Newton
    iteration per step (which needs two or more at 1e-8) run inside a mix of ordinary
    heat and two-phase requests.  Both come back flagged failed after their first step;
    every other request is bit-identical to its run alone; at the end no request is queued
    or in flight and the admitted footprint is back to zero.  The NaN is data, not a
    fault: the batch finds it in the first residual norm and retires the request before
    any Krylov solve (SURVEY's failure-isolation requirement; the reference's serial solve
    throws instead, EquelleRuntimeCPU.cpp:436-438)."""
    from equelle_amd.serving import twophase_request
    nan_field = gaussian_bumps(70, 70, 71, width_scale=0.04)
    nan_field[70 * 35 + 12] = np.nan
    reqs = [Request(0.0, 64, 0.2, 3, gaussian_bumps(64, 64, 70, width_scale=0.04)),
            Request(0.0, 70, 0.3, 3, nan_field),                       # NaN initial field
            twophase_request(0.0, 10, 2, seed=72),
            Request(0.0, 120, 0.45, 3, gaussian_bumps(120, 120, 73, width_scale=0.06)),   # Newton-limited
            Request(0.0, 45, 0.1, 2, gaussian_bumps(45, 45, 74, width_scale=0.03))]
    bad = {1, 3}
    srv = Server(hip_api, PARAMS + " hbm_budget=1e9")
    try:
        ids, results = {}, {}

        def tick():
            for rid in srv.progress():
                q = ids[rid]
                results[q] = srv.result(rid, reqs[q].n)

        for q, r in enumerate(reqs):
            rid = srv.submit(r)
            ids[rid] = q
            if q == 3:
                srv.set_max_newton(rid, 1)
            tick()   # later requests join mid-flight; the NaN request is retired in its first tick
        for _ in range(4000):
            if len(results) == len(reqs):
                break
            tick()
        assert len(results) == len(reqs)
        counts, used = srv.status()
        assert counts.tolist() == [0, 0] and used[0] == 0.0
    finally:
        srv.close()
    for q, r in enumerate(reqs):
        u, st, _ = results[q]
        if q in bad:
            assert st[3] == 1 and st[0] == 1, (q, st)   # failed after its first step
            continue
        assert st[3] == 0 and st[0] == r.steps, (q, st)
        assert np.array_equal(u, alone(hip_api, r)), q
