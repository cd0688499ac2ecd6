"""GPU parity at the configured workload sizes (the bench's own inputs and settings).

* nonlinear heat 2048 x 2048 (4,194,304 cells), one implicit step in the benchmark's mixed
  precision (FP32 Krylov on the red-black reduced system, FP64 refinement, Newton to 1e-8
  relative residual): the oracle's residual evaluated at the GPU's answer is below 1e-8 of
  its initial value, and the answer is within the distance that residual allows of the
  oracle's own step solved to 1e-12;
* two-phase 48 x 48 x 32 (73,728 cells): the multigrid V-cycle of the pressure stage runs
  its fine level and its 24 x 24 x 16 level as one-level kernels (k_mg_smooth, k_mg_resid,
  k_mg_restrict, k_mg_prolong_black) above the one-workgroup tail (12 x 12 x 8 and below);
  two steps in FP64 and in the benchmark's mixed precision against the oracle;
* the 512-instance batch of the batched workload (sides 32..512): every instance
  bit-identical to its run alone, and twelve sampled instances against the oracle.

Tolerance of a Newton answer: with J the Jacobian at the solution, ||u - u*|| <= ||J^-1||
||F(u)||; J = I + dt (cv vol)^-1 div k(u) grad (+ the k'(u) terms) has ||J^-1|| about 1 for
these implicit diffusion steps, and the tests allow a factor 10 on top.
"""
import ctypes as C

import numpy as np
import pytest

from equelle_amd.capi import dptr, iptr
from equelle_amd.workloads import NONLINEAR_HEAT, batched_instances, gaussian_bumps
from twophase_cases import case, oracle_run, twophase_params

pytestmark = pytest.mark.gpu

BENCH_HEAT = ("precision=mixed preconditioner=ilu0 rel_res_tol=1e-8 abs_res_tol=0 max_iter=20 "
              "solver_tol=1e-6 inner_tol=1e-6")


def oracle_heat_residual(api, nx, ny, k, u, u0, dt):
    """F(u; u0, dt) of the nonlinear heat program on the oracle (values, no Jacobian)."""
    n = nx * ny
    params = f"grid_dim=2 nx={nx} ny={ny} k={k} cv=1 k_model=quadratic".encode()
    e, ed = np.zeros(1, dtype=np.int32), np.zeros(1)
    F = np.zeros(n)
    api.heat_residual_values(params, dptr(u), dptr(u0), n, iptr(e), dptr(ed), 0, dt, dptr(F))
    return F


def oracle_heat_step(api, nx, ny, k, u0, dt, rel=1e-12, solver="auto"):
    params = (f"grid_dim=2 nx={nx} ny={ny} k={k} cv=1 k_model=quadratic abs_res_tol=0 rel_res_tol={rel} "
              f"max_iter=30 oracle_linear_solver={solver} oracle_solver_tol=1e-13").encode()
    e, ed = np.zeros(1, dtype=np.int32), np.zeros(1)
    out = np.zeros(nx * ny)
    it = np.zeros(1, dtype=np.int32)
    api.heat_run(params, dptr(u0), nx * ny, iptr(e), dptr(ed), 0, dptr(np.array([dt])), 1, dptr(out), iptr(it),
                 None)
    return out, int(it[0])


def batch(api, params):
    b = C.c_void_p()
    api.batch_create(params.encode(), C.byref(b))
    return b


def heat_add(api, b, nx, ny, k, u):
    s = C.c_int()
    api.batch_add(b, nx, ny, k, 1.0, dptr(np.ascontiguousarray(u)), C.byref(s))
    return s.value


def check_newton_answer(oracle_api, nx, ny, k, u0, dt, u, rel=1e-8):
    """The oracle's residual at u is below rel of its value at u0 (with 1% slack for the
    different summation order), and u is within 10 ||F(u)|| of the oracle's own answer."""
    f0 = np.linalg.norm(oracle_heat_residual(oracle_api, nx, ny, k, u0, u0, dt))
    fu = np.linalg.norm(oracle_heat_residual(oracle_api, nx, ny, k, u, u0, dt))
    assert fu <= 1.01 * rel * f0, (fu, f0)
    uo, _ = oracle_heat_step(oracle_api, nx, ny, k, u0, dt)
    fo = np.linalg.norm(oracle_heat_residual(oracle_api, nx, ny, k, uo, u0, dt))
    assert fo <= 1e-11 * f0, (fo, f0)
    err = np.linalg.norm(u - uo)
    assert err <= 10.0 * (fu + fo), (err, fu, fo)
    return err / np.linalg.norm(uo)


def test_heat_2048_mixed_step_matches_oracle(hip_api, oracle_api):
    nx = ny = NONLINEAR_HEAT["nx"]
    k, dt = NONLINEAR_HEAT["k"], NONLINEAR_HEAT["dt"]
    u0 = gaussian_bumps(nx, ny, seed=2048)
    print(f"heat {nx}x{ny}: fused mixed step", flush=True)
    b = batch(hip_api, BENCH_HEAT)
    try:
        s = heat_add(hip_api, b, nx, ny, k, u0)
        hip_api.batch_step(b, iptr(np.array([s], dtype=np.int32)), 1, dt)
        st, nrm = np.zeros(5, dtype=np.int32), np.zeros(2)
        hip_api.batch_stats(b, s, iptr(st), dptr(nrm))
        assert st[3] == 2 and st[4] == 0, (st, nrm)          # relative stop, not failed
        assert nrm[1] <= 1e-8 * nrm[0]
        u = np.zeros(nx * ny)
        hip_api.batch_state(b, s, dptr(u))
    finally:
        hip_api.batch_destroy(b)
    print("oracle residual and step at 2048^2", flush=True)
    rel = check_newton_answer(oracle_api, nx, ny, k, u0, dt, u)
    print(f"relative distance to the oracle's step: {rel:.3e}", flush=True)


@pytest.mark.parametrize("params,newton_tol,solver_tol,bar", [
    ("precision=fp64 preconditioner=ilu0", 1e-10, 1e-10, 1e-7),
    ("precision=mixed preconditioner=ilu0", 1e-8, 1e-6, 1e-5)])   # the benchmark's settings
def test_twophase_multigrid_above_tail_matches_oracle(hip_api, oracle_api, params, newton_tol, solver_tol, bar):
    nx, ny, nz = 48, 48, 32
    perm, poro, sw0, sc, sv = case(nx, ny, nz)
    n = perm.size
    dts = [1.0, 1.0]
    print(f"two-phase {nx}x{ny}x{nz}: oracle", flush=True)
    po, so, _ = oracle_run(oracle_api, twophase_params(nx, ny, nz, rel_res_tol=1e-11, max_iter=40,
                                                       extra="oracle_linear_solver=bicgstab "
                                                             "oracle_solver_tol=1e-13").encode(),
                           perm, poro, sw0, sc, sv, dts)
    print(f"two-phase {nx}x{ny}x{nz}: fused {params}", flush=True)
    b = batch(hip_api, params + f" rel_res_tol={newton_tol} abs_res_tol=0 solver_tol={solver_tol} max_iter=40 "
                                "solver_max_iter=5000")
    try:
        slot = C.c_int()
        hip_api.batch_add_twophase(b, b"watervisc=0.001 oilvisc=0.005 corey=2", nx, ny, nz, dptr(perm),
                                   dptr(poro), dptr(sw0), len(sc), iptr(sc), dptr(sv), C.byref(slot))
        st, tps = np.zeros(5, dtype=np.int32), np.zeros(3, dtype=np.int32)
        for dt in dts:
            hip_api.batch_step(b, iptr(np.array([slot.value], dtype=np.int32)), 1, dt)
            hip_api.batch_stats(b, slot.value, iptr(st), None)
            hip_api.batch_twophase_stats(b, slot.value, iptr(tps))
            assert st[3] in (1, 2) and tps[2] in (1, 2) and st[4] == 0, (st, tps)
        ph, sh = np.zeros(n), np.zeros(n)
        hip_api.batch_twophase_state(b, slot.value, dptr(ph), dptr(sh))
    finally:
        hip_api.batch_destroy(b)
    assert np.allclose(sh, so, rtol=bar, atol=0.1 * bar)
    assert np.allclose(ph - ph[-1], po - po[-1], rtol=bar, atol=bar * np.abs(po).max())


def test_batched_512_each_identical_alone_and_sampled_vs_oracle(hip_api, oracle_api):
    inst = batched_instances(512, seed=1234)
    dt = 1.0
    print(f"batched: {len(inst)} instances, {sum(s * s for s, _, _ in inst)} cells", flush=True)
    b = batch(hip_api, BENCH_HEAT)
    try:
        slots = np.array([heat_add(hip_api, b, s, s, k, u0) for s, k, u0 in inst], dtype=np.int32)
        hip_api.batch_step(b, iptr(slots), len(slots), dt)
        st = np.zeros((len(slots), 5), dtype=np.int32)
        hip_api.batch_stats_many(b, iptr(slots), len(slots), iptr(st))
        assert np.all(st[:, 3] == 2) and np.all(st[:, 4] == 0)
        together = []
        for (s, _, _), sl in zip(inst, slots):
            u = np.zeros(s * s)
            hip_api.batch_state(b, int(sl), dptr(u))
            together.append(u)
    finally:
        hip_api.batch_destroy(b)
    print("batched: each instance alone", flush=True)
    b = batch(hip_api, BENCH_HEAT)
    try:
        for q, (s, k, u0) in enumerate(inst):
            sl = heat_add(hip_api, b, s, s, k, u0)
            hip_api.batch_step(b, iptr(np.array([sl], dtype=np.int32)), 1, dt)
            u = np.zeros(s * s)
            hip_api.batch_state(b, sl, dptr(u))
            hip_api.batch_remove(b, sl)
            assert np.array_equal(u, together[q]), (q, s, np.abs(u - together[q]).max())
    finally:
        hip_api.batch_destroy(b)
    print("batched: twelve instances against the oracle", flush=True)
    rng = np.random.default_rng(12)
    order = np.argsort([s for s, _, _ in inst])
    sample = set(rng.choice(len(inst), 10, replace=False).tolist()) | {int(order[0]), int(order[-1])}
    for q in sorted(sample):
        s, k, u0 = inst[q]
        check_newton_answer(oracle_api, s, s, k, u0, dt, together[q])
