"""GPU parity of the fused heat engine (batched instances) against the oracle.

* residual + Jacobian of one assembly pass: CSR pattern bit-exact, values 1e-10 relative;
* implicit steps of several instances of different sizes and conductivities in one
  batch: each equals the oracle's newtonSolve within the tolerance the solves ran to;
* determinism: an instance's result is bit-identical alone and inside a batch;
* mixed precision (FP32 Krylov + FP64 refinement) converges to the FP64 answer.
"""
import ctypes as C

import numpy as np
import pytest

from equelle_amd.capi import dptr, iptr
from equelle_amd.workloads import gaussian_bumps

pytestmark = pytest.mark.gpu


def make_batch(api, params):
    b = C.c_void_p()
    api.batch_create(params.encode(), C.byref(b))
    return b


def add(api, b, nx, ny, k, u):
    slot = C.c_int()
    api.batch_add(b, nx, ny, k, 1.0, dptr(np.ascontiguousarray(u)), C.byref(slot))
    return slot.value


def state(api, b, slot, n):
    u = np.zeros(n)
    api.batch_state(b, slot, dptr(u))
    return u


def stats(api, b, slot):
    s = np.zeros(5, dtype=np.int32)
    nrm = np.zeros(2)
    api.batch_stats(b, slot, iptr(s), dptr(nrm))
    return s, nrm


def oracle_step(oracle_api, nx, ny, k, u0, dt, steps=1):
    params = (f"grid_dim=2 nx={nx} ny={ny} k={k} cv=1 k_model=quadratic abs_res_tol=0 rel_res_tol=1e-12 "
              "max_iter=30").encode()
    e, ed = np.zeros(1, dtype=np.int32), np.zeros(1)
    out = np.zeros(nx * ny)
    dts = np.full(steps, dt)
    oracle_api.heat_run(params, dptr(u0), nx * ny, iptr(e), dptr(ed), 0, dptr(dts), steps, dptr(out), None, None)
    return out


@pytest.mark.parametrize("nx,ny", [(64, 48), (257, 129)])
def test_fused_residual_matches_oracle(hip_api, oracle_api, nx, ny):
    b = make_batch(hip_api, "precision=fp64 preconditioner=ilu0")
    try:
        rng = np.random.default_rng(nx)
        u, u0 = rng.uniform(0, 2, nx * ny), rng.uniform(0, 2, nx * ny)
        slot = add(hip_api, b, nx, ny, 0.13, u)
        n = nx * ny
        info = np.zeros(2, dtype=np.int32)
        F = np.zeros(n)
        rp = np.zeros(n + 1, dtype=np.int32)
        ci = np.zeros(5 * n, dtype=np.int32)
        v = np.zeros(5 * n)
        hip_api.batch_residual(b, slot, dptr(u), dptr(u0), 0.7, iptr(info), dptr(F), iptr(rp), iptr(ci), dptr(v))
        nnz = info[1]
        params = f"grid_dim=2 nx={nx} ny={ny} k=0.13 cv=1 k_model=quadratic".encode()
        oi = np.zeros(2, dtype=np.int32)
        e, ed = np.zeros(1, dtype=np.int32), np.zeros(1)
        oracle_api.heat_residual(params, dptr(u), dptr(u0), n, iptr(e), dptr(ed), 0, 0.7, iptr(oi), None, None, None,
                                 None)
        Fo = np.zeros(n)
        rpo = np.zeros(n + 1, dtype=np.int32)
        cio = np.zeros(oi[1], dtype=np.int32)
        vo = np.zeros(oi[1])
        oracle_api.heat_residual(params, dptr(u), dptr(u0), n, iptr(e), dptr(ed), 0, 0.7, iptr(oi), dptr(Fo),
                                 iptr(rpo), iptr(cio), dptr(vo))
        assert nnz == oi[1] and np.array_equal(rp, rpo) and np.array_equal(ci[:nnz], cio)
        assert np.allclose(F, Fo, rtol=1e-10, atol=1e-13) and np.allclose(v[:nnz], vo, rtol=1e-10, atol=1e-13)
    finally:
        hip_api.batch_destroy(b)


@pytest.mark.parametrize("params", ["precision=fp64 preconditioner=ilu0",
                                    "precision=fp64 preconditioner=diagonal",
                                    "precision=mixed preconditioner=ilu0",
                                    "precision=mixed preconditioner=diagonal",
                                    "precision=mixed preconditioner=none"])
def test_fused_batch_steps_match_oracle(hip_api, oracle_api, params):
    b = make_batch(hip_api, params + " rel_res_tol=1e-12 abs_res_tol=0 solver_tol=1e-12 max_iter=30 "
                                     "solver_max_iter=3000")
    try:
        sizes = [(40, 33, 0.1), (97, 64, 0.45), (31, 128, 0.05), (128, 128, 0.2)]
        slots, u0s = [], []
        for q, (nx, ny, k) in enumerate(sizes):
            u0 = gaussian_bumps(nx, ny, 100 + q, width_scale=max(nx, ny) / 256)
            u0s.append(u0)
            slots.append(add(hip_api, b, nx, ny, k, u0))
        arr = np.array(slots, dtype=np.int32)
        hip_api.batch_step(b, iptr(arr), len(arr), 1.0)
        hip_api.batch_step(b, iptr(arr), len(arr), 2.0)
        for (nx, ny, k), slot, u0 in zip(sizes, slots, u0s):
            s, nrm = stats(hip_api, b, slot)
            assert s[3] in (1, 2) and s[4] == 0, (s, nrm)
            got = state(hip_api, b, slot, nx * ny)
            want = oracle_step(oracle_api, nx, ny, k, oracle_step(oracle_api, nx, ny, k, u0, 1.0), 2.0)
            assert np.allclose(got, want, rtol=1e-9, atol=1e-11), np.abs(got - want).max()
    finally:
        hip_api.batch_destroy(b)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("nx", [77, 80])   # 80: four-ranks-per-lane Krylov kernels
def test_result_identical_alone_and_in_batch(hip_api, precision, nx):
    params = f"precision={precision} preconditioner=ilu0 rel_res_tol=1e-8 solver_tol=1e-6"
    ny, k = 53, 0.3
    u0 = gaussian_bumps(nx, ny, 7, width_scale=0.05)
    alone = make_batch(hip_api, params)
    mixed = make_batch(hip_api, params)
    try:
        s = add(hip_api, alone, nx, ny, k, u0)
        hip_api.batch_step(alone, iptr(np.array([s], dtype=np.int32)), 1, 1.0)
        ua = state(hip_api, alone, s, nx * ny)
        others = [add(hip_api, mixed, 150, 90, 0.2, gaussian_bumps(150, 90, 8, width_scale=0.1)),
                  add(hip_api, mixed, 33, 33, 0.07, gaussian_bumps(33, 33, 9, width_scale=0.02))]
        s2 = add(hip_api, mixed, nx, ny, k, u0)
        others.append(add(hip_api, mixed, 300, 17, 0.5, gaussian_bumps(300, 17, 10, width_scale=0.1)))
        arr = np.array(others[:2] + [s2] + others[2:], dtype=np.int32)
        hip_api.batch_step(mixed, iptr(arr), len(arr), 1.0)
        ub = state(hip_api, mixed, s2, nx * ny)
        assert np.array_equal(ua, ub)
        assert np.array_equal(stats(hip_api, alone, s)[0], stats(hip_api, mixed, s2)[0])
    finally:
        hip_api.batch_destroy(alone)
        hip_api.batch_destroy(mixed)


def test_continuous_batching_join_mid_flight(hip_api):
    """A slot that starts its step while another is mid-Newton gets the same result."""
    params = "precision=mixed preconditioner=ilu0 rel_res_tol=1e-8 solver_tol=1e-6"
    b = make_batch(hip_api, params)
    ref = make_batch(hip_api, params)
    try:
        ua = gaussian_bumps(200, 150, 11, width_scale=0.1)
        ub = gaussian_bumps(64, 80, 12, width_scale=0.05)
        sa = add(hip_api, b, 200, 150, 0.3, ua)
        sb = add(hip_api, b, 64, 80, 0.1, ub)
        hip_api.batch_begin_step(b, sa, 1.0)
        fin = np.zeros(8, dtype=np.int32)
        nf = C.c_int()
        hip_api.batch_iterate(b, iptr(fin), 8, C.byref(nf))   # sa: assembled, in flight
        hip_api.batch_begin_step(b, sb, 1.0)                  # sb joins the running batch
        flag = C.c_int(1)
        guard = 0
        while True:
            hip_api.batch_iterate(b, iptr(fin), 8, C.byref(nf))
            fa, fb = C.c_int(), C.c_int()
            hip_api.batch_in_flight(b, sa, C.byref(fa))
            hip_api.batch_in_flight(b, sb, C.byref(fb))
            if not fa.value and not fb.value:
                break
            guard += 1
            assert guard < 100
        rb = add(hip_api, ref, 64, 80, 0.1, ub)
        hip_api.batch_step(ref, iptr(np.array([rb], dtype=np.int32)), 1, 1.0)
        assert np.array_equal(state(hip_api, b, sb, 64 * 80), state(hip_api, ref, rb, 64 * 80))
    finally:
        hip_api.batch_destroy(b)
        hip_api.batch_destroy(ref)


def _run_to_done(api, b, slots):
    fin = np.zeros(8, dtype=np.int32)
    nf = C.c_int()
    for _ in range(200):
        api.batch_iterate(b, iptr(fin), 8, C.byref(nf))
        flags = []
        for s in slots:
            f = C.c_int()
            api.batch_in_flight(b, s, C.byref(f))
            flags.append(f.value)
        if not any(flags):
            return
    raise AssertionError("batch did not finish")


def test_step_start_copies_before_remove_and_set_state(hip_api):
    """begin_step queues each instance's step-start copy u0 = u, and the queued copies are
    issued together before the next device work.  Removing a slot while copies are queued
    leaves the others stepping exactly as a plain batch step; set_state after begin_step
    keeps u0 at the state the step began from (the step solves the same implicit equation
    from another Newton start)."""
    params = "precision=mixed preconditioner=ilu0 rel_res_tol=1e-10 solver_tol=1e-6"
    dims = [(64, 48, 0.3), (40, 56, 0.2), (72, 24, 0.15)]
    us = [gaussian_bumps(nx, ny, 21 + q, width_scale=0.1) for q, (nx, ny, _) in enumerate(dims)]
    b, ref = make_batch(hip_api, params), make_batch(hip_api, params)
    try:
        slots = [add(hip_api, b, nx, ny, k, u) for (nx, ny, k), u in zip(dims, us)]
        for s in slots:
            hip_api.batch_begin_step(b, s, 1.0)
        hip_api.batch_remove(b, slots[0])   # three copies queued, one of them the removed slot's
        _run_to_done(hip_api, b, slots[1:])
        rs = [add(hip_api, ref, nx, ny, k, u) for (nx, ny, k), u in zip(dims[1:], us[1:])]
        hip_api.batch_step(ref, iptr(np.array(rs, dtype=np.int32)), len(rs), 1.0)
        for s, r, (nx, ny, _) in zip(slots[1:], rs, dims[1:]):
            assert np.array_equal(state(hip_api, b, s, nx * ny), state(hip_api, ref, r, nx * ny))
    finally:
        hip_api.batch_destroy(b)
        hip_api.batch_destroy(ref)
    nx, ny, k = dims[0]
    b, ref = make_batch(hip_api, params), make_batch(hip_api, params)
    try:
        s = add(hip_api, b, nx, ny, k, us[0])
        hip_api.batch_begin_step(b, s, 1.0)
        start = np.ascontiguousarray(1.1 * us[0] + 0.01)
        hip_api.batch_set_state(b, s, dptr(start))   # the Newton start changes, u0 does not
        _run_to_done(hip_api, b, [s])
        r = add(hip_api, ref, nx, ny, k, us[0])
        hip_api.batch_step(ref, iptr(np.array([r], dtype=np.int32)), 1, 1.0)
        got, want = state(hip_api, b, s, nx * ny), state(hip_api, ref, r, nx * ny)
        # u0 taken after set_state would solve the step from `start`: a 10% different field
        assert np.allclose(got, want, rtol=1e-7, atol=1e-9 * np.abs(want).max())
    finally:
        hip_api.batch_destroy(b)
        hip_api.batch_destroy(ref)


def test_stats_many_matches_stats(hip_api):
    b = make_batch(hip_api, "precision=mixed preconditioner=ilu0 rel_res_tol=1e-8 solver_tol=1e-6")
    try:
        dims = [(48, 40, 0.2), (64, 32, 0.4), (30, 30, 0.1)]
        slots = np.array([add(hip_api, b, nx, ny, k, gaussian_bumps(nx, ny, 31 + q, width_scale=0.1))
                          for q, (nx, ny, k) in enumerate(dims)], dtype=np.int32)
        hip_api.batch_step(b, iptr(slots), len(slots), 1.0)
        many = np.zeros((len(slots), 5), dtype=np.int32)
        hip_api.batch_stats_many(b, iptr(slots), len(slots), iptr(many))
        for q, sl in enumerate(slots):
            assert many[q].tolist() == stats(hip_api, b, int(sl))[0].tolist()
        assert (many[:, 0] > 0).all() and (many[:, 3] > 0).all()
    finally:
        hip_api.batch_destroy(b)


def test_mixed_matches_fp64_on_large_grid(hip_api):
    nx = ny = 512
    u0 = gaussian_bumps(nx, ny, 2048, width_scale=nx / 2048)
    out = {}
    for prec in ("fp64", "mixed"):
        b = make_batch(hip_api, f"precision={prec} preconditioner=ilu0 rel_res_tol=1e-8 solver_tol=1e-6")
        try:
            s = add(hip_api, b, nx, ny, 0.1, u0)
            hip_api.batch_step(b, iptr(np.array([s], dtype=np.int32)), 1, 1.0)
            st, nrm = stats(hip_api, b, s)
            assert st[3] == 2 and nrm[1] <= 1e-8 * nrm[0]
            out[prec] = state(hip_api, b, s, nx * ny)
        finally:
            hip_api.batch_destroy(b)
    assert np.linalg.norm(out["mixed"] - out["fp64"]) <= 1e-7 * np.linalg.norm(out["fp64"])


def test_graph_replays_match_eager(hip_api):
    """The BiCGStab iterations captured in a hipGraph and replayed give bit-identical
    results to eager launches, over several steps (every replay checked)."""
    nx, ny = 96, 72
    u0 = gaussian_bumps(nx, ny, 21, width_scale=0.05)
    res = {}
    for graph in (0, 1):
        b = make_batch(hip_api, f"precision=mixed preconditioner=ilu0 rel_res_tol=1e-8 solver_tol=1e-6 graph={graph}")
        try:
            s = add(hip_api, b, nx, ny, 0.25, u0)
            traj = []
            for dt in (0.5, 1.0, 1.0, 2.0):
                hip_api.batch_step(b, iptr(np.array([s], dtype=np.int32)), 1, dt)
                traj.append(state(hip_api, b, s, nx * ny))
            res[graph] = traj
        finally:
            hip_api.batch_destroy(b)
    for a, g in zip(res[0], res[1]):
        assert np.array_equal(a, g)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_speculative_solves_match_checked_solves(hip_api, precision):
    """Queuing the work after a Krylov solve behind the predicted iterations (gated on the
    device by the convergence count) gives the same results as waiting for the count.  The
    state is replaced by a rough field between steps, so the iteration predictions of the
    smooth steps fall short and the gated work is queued again after the catch-up."""
    sizes = [(96, 72), (64, 64)]
    rng = np.random.default_rng(5)
    smooth = [1.0 + 0.01 * gaussian_bumps(nx, ny, 3 + q, width_scale=0.2) for q, (nx, ny) in enumerate(sizes)]
    rough = [rng.uniform(0.0, 2.0, nx * ny) for nx, ny in sizes]
    res = {}
    for spec in (0, 1):
        b = make_batch(hip_api, f"precision={precision} preconditioner=ilu0 rel_res_tol=1e-8 solver_tol=1e-6 "
                                f"graph=1 speculate={spec}")
        try:
            slots = np.array([add(hip_api, b, nx, ny, 0.3, u) for (nx, ny), u in zip(sizes, smooth)],
                             dtype=np.int32)
            traj = []
            for phase, fields in (("smooth", None), ("rough", rough)):
                if fields is not None:
                    for sl, u in zip(slots, fields):
                        hip_api.batch_set_state(b, int(sl), dptr(u))
                for _ in range(3):
                    hip_api.batch_step(b, iptr(slots), len(slots), 1.0)
                    for sl, (nx, ny) in zip(slots, sizes):
                        st, _ = stats(hip_api, b, int(sl))
                        assert st[3] != 0 and st[4] == 0, (phase, st)
                        traj.append((state(hip_api, b, int(sl), nx * ny), st[:3].tolist()))
            res[spec] = traj
        finally:
            hip_api.batch_destroy(b)
    for (ua, sa), (ub, sb) in zip(res[0], res[1]):
        assert sa == sb
        assert np.array_equal(ua, ub)
    # the rough steps need more Krylov iterations than the smooth ones predicted
    half = len(res[1]) // 2
    assert sum(s[1] for _, s in res[1][half:]) > sum(s[1] for _, s in res[1][:half])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_predicted_last_newton_check_matches_full_assembly(hip_api, precision):
    """After a Newton iteration that the instance's previous step says is the last, the
    assembly forms ||F|| alone.  A rough field after smooth steps needs more iterations
    than predicted: the full assembly then follows, and the step is bit-identical to the
    same step taken by a fresh batch (no prediction, every assembly full).  forcing0=0: the
    first linear tolerance of a step does not depend on the previous step either."""
    nx, ny = 96, 80
    rng = np.random.default_rng(11)
    smooth = 1.0 + 0.01 * gaussian_bumps(nx, ny, 7, width_scale=0.2)
    rough = rng.uniform(0.0, 2.0, nx * ny)
    params = (f"precision={precision} preconditioner=ilu0 rel_res_tol=1e-9 abs_res_tol=0 solver_tol=1e-6 "
              "forcing0=0")
    b = make_batch(hip_api, params)
    try:
        s = add(hip_api, b, nx, ny, 0.3, smooth)
        arr = np.array([s], dtype=np.int32)
        hip_api.batch_step(b, iptr(arr), 1, 1.0)
        hip_api.batch_step(b, iptr(arr), 1, 1.0)
        n_smooth = stats(hip_api, b, s)[0][0]
        hip_api.batch_set_state(b, s, dptr(rough))
        hip_api.batch_step(b, iptr(arr), 1, 1.0)
        st_a, nrm_a = stats(hip_api, b, s)
        ua = state(hip_api, b, s, nx * ny)
    finally:
        hip_api.batch_destroy(b)
    assert st_a[0] > n_smooth, (st_a, n_smooth)   # the prediction fell short
    b = make_batch(hip_api, params)
    try:
        s = add(hip_api, b, nx, ny, 0.3, rough)
        hip_api.batch_step(b, iptr(np.array([s], dtype=np.int32)), 1, 1.0)
        st_b, nrm_b = stats(hip_api, b, s)
        ub = state(hip_api, b, s, nx * ny)
    finally:
        hip_api.batch_destroy(b)
    assert st_a[0] == st_b[0] and st_a[3] != 0, (st_a, st_b)
    assert np.array_equal(nrm_a, nrm_b)
    assert np.array_equal(ua, ub)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_converged_step_start_leaves_state_unchanged(hip_api, precision):
    """A step start whose ||F_0|| already meets abs_res_tol: with the previous steps' prediction
    the first assembly and the first solve run as one graph before the host sees ||F_0||; the
    solve must then be a no-op (no Newton iteration counted, u unchanged, converged)."""
    nx, ny = 96, 64
    params = f"precision={precision} preconditioner=ilu0 rel_res_tol=1e-9 abs_res_tol=1e-12 solver_tol=1e-6"
    b = make_batch(hip_api, params)
    try:
        s = add(hip_api, b, nx, ny, 0.2, gaussian_bumps(nx, ny, 3, width_scale=0.1))
        arr = np.array([s], dtype=np.int32)
        for _ in range(2):
            hip_api.batch_step(b, iptr(arr), 1, 1.0)
        assert stats(hip_api, b, s)[0][0] >= 1
        flat = np.full(nx * ny, 0.7)   # uniform: no flux, F_0 = 0
        hip_api.batch_set_state(b, s, dptr(flat))
        hip_api.batch_step(b, iptr(arr), 1, 1.0)
        st, nrm = stats(hip_api, b, s)
        assert st[0] == 0 and st[3] == 1 and st[4] == 0, (st, nrm)
        assert np.array_equal(state(hip_api, b, s, nx * ny), flat)
        # and the next non-trivial step runs normally
        hip_api.batch_set_state(b, s, dptr(gaussian_bumps(nx, ny, 4, width_scale=0.1)))
        hip_api.batch_step(b, iptr(arr), 1, 1.0)
        st, nrm = stats(hip_api, b, s)
        assert st[0] >= 1 and st[3] != 0 and st[4] == 0, (st, nrm)
    finally:
        hip_api.batch_destroy(b)


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("mixed", 1e-9)])
def test_pair_vectors_match_strided_form(hip_api, precision, tol):
    """The four-ranks-per-lane Krylov kernels (2D heat, nx % 8 == 0) and the strided
    ones solve the same systems: the steps agree to the solve tolerance (the partial
    sums are added in another order, so the iterates differ in rounding)."""
    sizes = [(256, 192, 0.2), (512, 512, 0.1), (64, 8, 0.4)]
    out = {}
    for pv in (0, 1):
        b = make_batch(hip_api, f"precision={precision} preconditioner=ilu0 rel_res_tol=1e-10 abs_res_tol=0 "
                                f"solver_tol=1e-10 max_iter=30 pair_vectors={pv}")
        try:
            slots = [add(hip_api, b, nx, ny, k, gaussian_bumps(nx, ny, 40 + q, width_scale=nx / 2048))
                     for q, (nx, ny, k) in enumerate(sizes)]
            arr = np.array(slots, dtype=np.int32)
            hip_api.batch_step(b, iptr(arr), len(arr), 1.0)
            hip_api.batch_step(b, iptr(arr), len(arr), 1.0)
            res = []
            for s, (nx, ny, _) in zip(slots, sizes):
                st, nrm = stats(hip_api, b, s)
                assert st[3] in (1, 2) and st[4] == 0, (pv, st, nrm)
                res.append(state(hip_api, b, s, nx * ny))
            out[pv] = res
        finally:
            hip_api.batch_destroy(b)
    for a, v in zip(out[0], out[1]):
        assert np.linalg.norm(a - v) <= tol * np.linalg.norm(a), np.linalg.norm(a - v) / np.linalg.norm(a)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_operator_tiles_match_one_pass_operator(hip_api, precision):
    """The reduced operator in tiles of four rows x 256 pair ranks (nx % 512 == 0, red
    unknowns shared through LDS) gives steps bit-identical to the one-pass operator, next
    to instances that take the one-pass operator (nx = 80) and the strided kernels (nx = 70)
    in the same batch; the heights are not multiples of four (tiles below the grid)."""
    sizes = [(512, 94, 0.2), (1024, 37, 0.1), (80, 64, 0.3), (70, 45, 0.4)]
    out = {}
    for tiles in (0, 1):
        b = make_batch(hip_api, f"precision={precision} preconditioner=ilu0 rel_res_tol=1e-9 abs_res_tol=0 "
                                f"solver_tol=1e-6 operator_tiles={tiles}")
        try:
            slots = [add(hip_api, b, nx, ny, k, gaussian_bumps(nx, ny, 50 + q, width_scale=0.05))
                     for q, (nx, ny, k) in enumerate(sizes)]
            arr = np.array(slots, dtype=np.int32)
            res = []
            for _ in range(2):
                hip_api.batch_step(b, iptr(arr), len(arr), 1.0)
                for sl, (nx, ny, _) in zip(slots, sizes):
                    st, _ = stats(hip_api, b, sl)
                    assert st[3] != 0 and st[4] == 0, (tiles, st)
                    res.append((state(hip_api, b, sl, nx * ny), st[:3].tolist()))
            out[tiles] = res
        finally:
            hip_api.batch_destroy(b)
    for (ua, sa), (ub, sb) in zip(out[0], out[1]):
        assert sa == sb
        assert np.array_equal(ua, ub)


def test_launch_shapes_give_identical_results(hip_api):
    """Every launch shape (cells per thread 1/2/4/8, XCD block order on/off) gives
    bit-identical steps: reductions use fixed groups of cells, not blocks."""
    sizes = [(70, 45, 0.2), (33, 130, 0.4), (256, 200, 0.1)]
    ref = None
    for cpt in (1, 2, 4, 8):
        for xcd in (0, 1):
            b = make_batch(hip_api, f"precision=mixed preconditioner=ilu0 rel_res_tol=1e-8 solver_tol=1e-6 "
                                    f"cpt={cpt} xcd={xcd}")
            try:
                slots = [add(hip_api, b, nx, ny, k, gaussian_bumps(nx, ny, 30 + q, width_scale=0.05))
                         for q, (nx, ny, k) in enumerate(sizes)]
                hip_api.batch_step(b, iptr(np.array(slots, dtype=np.int32)), len(slots), 1.0)
                info = np.zeros(3, dtype=np.int32)
                hip_api.batch_shape(b, iptr(info))
                assert info[0] == cpt and info[1] == xcd
                out = [state(hip_api, b, s, nx * ny) for s, (nx, ny, _) in zip(slots, sizes)]
            finally:
                hip_api.batch_destroy(b)
            if ref is None:
                ref = out
            else:
                for a, c in zip(ref, out):
                    assert np.array_equal(a, c), (cpt, xcd)


def test_tuner_picks_a_shape(hip_api):
    b = make_batch(hip_api, "precision=mixed preconditioner=ilu0 rel_res_tol=1e-8 solver_tol=1e-6")
    try:
        s = add(hip_api, b, 300, 200, 0.2, gaussian_bumps(300, 200, 5, width_scale=0.1))
        hip_api.batch_step(b, iptr(np.array([s], dtype=np.int32)), 1, 1.0)
        info = np.zeros(3, dtype=np.int32)
        hip_api.batch_shape(b, iptr(info))
        assert info[0] in (1, 2, 4, 8) and info[1] in (0, 1) and info[2] >= 1
        st, nrm = stats(hip_api, b, s)
        assert st[3] == 2 and nrm[1] <= 1e-8 * nrm[0]
    finally:
        hip_api.batch_destroy(b)
