"""GPU parity of the two-phase program (examples/dsl/twophase.equelle with Corey
exponent 2) on the device AD engine against the oracle.

* pressure and transport residual + Jacobian: CSR pattern bit-exact, values 1e-10;
* time steps (pressure newtonSolve + saturation newtonSolve): saturation within
  1e-9 of the oracle's (GMRES; 1e-7 for BiCGStab), pressure within the same bar up to
  the additive constant the no-flow pressure equation leaves free (compared after
  subtracting the last cell's value);
* FP32 GMRES(30) + Jacobi with FP64 refinement converges to the FP64 result;
* the fully implicit program (twophase_fully_implicit.equelle, newtonSolveSystem): the
  combined 2n x 2n residual + Jacobian bit-exact in pattern and 1e-10 in values, and its
  steps within 1e-9 of the oracle's (GMRES with Jacobi or ILU(0)).
"""
import ctypes as C

import numpy as np
import pytest

from equelle_amd.capi import dptr, iptr
from twophase_cases import (case, fi_params, oracle_fi_residual, oracle_fi_run, oracle_residual, oracle_run,
                            twophase_params)

pytestmark = pytest.mark.gpu


def create(api, params, perm, poro, sw0, sc, sv):
    sim = C.c_void_p()
    api.twophase_create(params, dptr(perm), dptr(poro), dptr(sw0), perm.size, iptr(sc), dptr(sv), len(sc),
                        C.byref(sim))
    return sim


def hip_residual(api, sim, which, p, sw0, sw, dt, n):
    info = np.zeros(2, dtype=np.int32)
    api.twophase_residual(sim, which, dptr(p), dptr(sw0), dptr(sw), dt, iptr(info), None, None, None, None)
    F, rp = np.zeros(n), np.zeros(n + 1, dtype=np.int32)
    ci, v = np.zeros(info[1], dtype=np.int32), np.zeros(info[1])
    api.twophase_residual(sim, which, dptr(p), dptr(sw0), dptr(sw), dt, iptr(info), dptr(F), iptr(rp), iptr(ci),
                          dptr(v))
    return F, rp, ci, v


@pytest.mark.parametrize("dims", [(9, 7, 5), (16, 12, 3)])
@pytest.mark.parametrize("which", [0, 1])
def test_twophase_residual_jacobian_parity(hip_api, oracle_api, dims, which):
    nx, ny, nz = dims
    perm, poro, sw0, sc, sv = case(nx, ny, nz, seed=nx)
    n = perm.size
    params = twophase_params(nx, ny, nz).encode()
    rng = np.random.default_rng(nz)
    p, s0, sw = rng.uniform(0, 1, n), rng.uniform(0, 1, n), rng.uniform(0.05, 0.95, n)
    sim = create(hip_api, params, perm, poro, sw0, sc, sv)
    try:
        Fh, rph, cih, vh = hip_residual(hip_api, sim, which, p, s0, sw, 0.8, n)
    finally:
        hip_api.twophase_destroy(sim)
    Fo, rpo, cio, vo = oracle_residual(oracle_api, params, perm, poro, sc, sv, which, p, s0, sw, 0.8)
    assert np.array_equal(rph, rpo) and np.array_equal(cih, cio)
    scale = np.abs(Fo).max()
    assert np.allclose(Fh, Fo, rtol=1e-10, atol=1e-12 * scale)
    assert np.allclose(vh, vo, rtol=1e-10, atol=1e-12 * np.abs(vo).max())


# BiCGStab cannot drive the singular (no-flow) pressure system's residual as far below
# rounding as GMRES, so its case runs to looser tolerances.
@pytest.mark.parametrize("extra,newton_tol,solver_tol,bar", [
    ("precision=fp64 solver=GMRes preconditioner=diagonal", 1e-12, 1e-13, 1e-9),
    ("precision=fp64 solver=BiCGStab preconditioner=diagonal", 1e-10, 1e-11, 1e-7),
    ("precision=mixed solver=GMRes preconditioner=diagonal", 1e-12, 1e-13, 1e-9)])
def test_twophase_steps_match_oracle(hip_api, oracle_api, extra, newton_tol, solver_tol, bar):
    nx, ny, nz = 12, 10, 6
    perm, poro, sw0, sc, sv = case(nx, ny, nz)
    n = perm.size
    dts = [1.0, 1.0, 2.0]
    oparams = twophase_params(nx, ny, nz, rel_res_tol=1e-13, max_iter=40).encode()
    po, so, _ = oracle_run(oracle_api, oparams, perm, poro, sw0, sc, sv, dts)
    hparams = twophase_params(nx, ny, nz, rel_res_tol=newton_tol, max_iter=40, solver_tol=solver_tol,
                              extra=extra).encode()
    sim = create(hip_api, hparams, perm, poro, sw0, sc, sv)
    try:
        st = np.zeros(6, dtype=np.int32)
        for dt in dts:
            hip_api.twophase_step(sim, dt, iptr(st), None)
            assert st[2] in (1, 2) and st[5] in (1, 2), st
        ph, sh = np.zeros(n), np.zeros(n)
        hip_api.twophase_state(sim, dptr(ph), dptr(sh))
    finally:
        hip_api.twophase_destroy(sim)
    assert np.allclose(sh, so, rtol=bar, atol=0.1 * bar)
    dp_h, dp_o = ph - ph[-1], po - po[-1]
    assert np.allclose(dp_h, dp_o, rtol=bar, atol=bar * np.abs(dp_o).max())


# ---------------------------------------------------------------- fully implicit program
# examples/dsl/twophase_fully_implicit.equelle: newtonSolveSystem on (p, sw) through the
# facade (subset / superset of the combined primary variable, 2n x 2n Jacobian with the
# coupling blocks).
def hip_fi_residual(api, sim, p, sw0, sw, dt, n):
    info = np.zeros(2, dtype=np.int32)
    api.twophase_residual(sim, 2, dptr(p), dptr(sw0), dptr(sw), dt, iptr(info), None, None, None, None)
    F, rp = np.zeros(info[0]), np.zeros(info[0] + 1, dtype=np.int32)
    ci, v = np.zeros(info[1], dtype=np.int32), np.zeros(info[1])
    api.twophase_residual(sim, 2, dptr(p), dptr(sw0), dptr(sw), dt, iptr(info), dptr(F), iptr(rp), iptr(ci),
                          dptr(v))
    return F, rp, ci, v


@pytest.mark.parametrize("dims", [(6, 5, 4), (11, 7, 3)])
def test_fully_implicit_residual_jacobian_parity(hip_api, oracle_api, dims):
    nx, ny, nz = dims
    perm, poro, sw0, sc, sv = case(nx, ny, nz, seed=nx + 1)
    n = perm.size
    params = fi_params(nx, ny, nz, extra="fully_implicit=1").encode()
    rng = np.random.default_rng(nx)
    p, s0, sw = rng.uniform(0, 1, n), rng.uniform(0, 1, n), rng.uniform(0.05, 0.95, n)
    sim = create(hip_api, params, perm, poro, sw0, sc, sv)
    try:
        Fh, rph, cih, vh = hip_fi_residual(hip_api, sim, p, s0, sw, 0.3, n)
    finally:
        hip_api.twophase_destroy(sim)
    Fo, rpo, cio, vo = oracle_fi_residual(oracle_api, params, perm, poro, sc, sv, p, s0, sw, 0.3)
    assert Fh.size == 2 * n
    assert np.array_equal(rph, rpo) and np.array_equal(cih, cio)
    assert np.allclose(Fh, Fo, rtol=1e-10, atol=1e-12 * np.abs(Fo).max())
    assert np.allclose(vh, vo, rtol=1e-10, atol=1e-12 * np.abs(vo).max())


# GMRES: the coupled system is singular (pressure constants); BiCGStab stagnates at its
# rounding-inconsistency floor there and Newton then stalls above 1e-12.
@pytest.mark.parametrize("extra,bar", [
    ("precision=fp64 solver=GMRes preconditioner=diagonal", 1e-9),
    ("precision=fp64 solver=GMRes preconditioner=ILU0", 1e-9),
    ("precision=fp64 solver=GMRes preconditioner=cpr", 1e-9)])
def test_fully_implicit_steps_match_oracle(hip_api, oracle_api, extra, bar):
    nx, ny, nz = 8, 7, 4
    perm, poro, sw0, sc, sv = case(nx, ny, nz)
    n = perm.size
    # plain Newton from the program's initial guess (p0, 0.5) diverges for larger steps
    # on this case (the oracle needs 40+ iterations at dt 0.15); these converge
    dts = [0.05, 0.05, 0.1]
    oparams = fi_params(nx, ny, nz, rel_res_tol=1e-13, max_iter=40).encode()
    po, so, oit = oracle_fi_run(oracle_api, oparams, perm, poro, sw0, sc, sv, dts)
    assert oit.max() < 40, oit
    # inexact Newton: the coupled system is singular (pressure constants), so its rounding
    # inconsistency floors the linear residual; 1e-7 keeps the solves above that floor
    # while Newton still converges quadratically to the 1e-12 residual
    hparams = fi_params(nx, ny, nz, rel_res_tol=1e-12, max_iter=40, solver_tol=1e-7,
                        extra=extra + " fully_implicit=1").encode()
    sim = create(hip_api, hparams, perm, poro, sw0, sc, sv)
    try:
        st = np.zeros(6, dtype=np.int32)
        for dt in dts:
            hip_api.twophase_step(sim, dt, iptr(st), None)
            assert st[2] in (1, 2) and st[5] in (1, 2) and st[0] >= 2, st
        ph, sh = np.zeros(n), np.zeros(n)
        hip_api.twophase_state(sim, dptr(ph), dptr(sh))
    finally:
        hip_api.twophase_destroy(sim)
    assert np.allclose(sh, so, rtol=bar, atol=0.1 * bar)
    dp_h, dp_o = ph - ph[-1], po - po[-1]
    assert np.allclose(dp_h, dp_o, rtol=bar, atol=bar * np.abs(dp_o).max())


# ---------------------------------------------------------------- fused batch engine
def fused_batch(api, params):
    b = C.c_void_p()
    api.batch_create(params.encode(), C.byref(b))
    return b


TP_INST = "watervisc=0.001 oilvisc=0.005 corey=2"


def fused_add(api, b, nx, ny, nz, perm, poro, sw0, sc, sv):
    slot = C.c_int()
    api.batch_add_twophase(b, TP_INST.encode(), nx, ny, nz, dptr(perm), dptr(poro), dptr(sw0), len(sc), iptr(sc),
                           dptr(sv), C.byref(slot))
    return slot.value


@pytest.mark.parametrize("dims", [(9, 7, 5), (16, 12, 3), (11, 6, 1)])
@pytest.mark.parametrize("stage", [0, 1])
def test_fused_twophase_residual_matches_oracle(hip_api, oracle_api, dims, stage):
    nx, ny, nz = dims
    perm, poro, sw0, sc, sv = case(nx, ny, nz, seed=nx + nz)
    n = perm.size
    rng = np.random.default_rng(3 * nx + stage)
    p, s0, sw = rng.uniform(0, 1, n), rng.uniform(0, 1, n), rng.uniform(0.05, 0.95, n)
    b = fused_batch(hip_api, "precision=fp64 preconditioner=ilu0")
    try:
        slot = fused_add(hip_api, b, nx, ny, nz, perm, poro, sw0, sc, sv)
        info = np.zeros(2, dtype=np.int32)
        F, rp = np.zeros(n), np.zeros(n + 1, dtype=np.int32)
        ci, v = np.zeros(7 * n, dtype=np.int32), np.zeros(7 * n)
        hip_api.batch_twophase_residual(b, slot, stage, dptr(p), dptr(s0), dptr(sw), 0.8, iptr(info), dptr(F),
                                        iptr(rp), iptr(ci), dptr(v))
    finally:
        hip_api.batch_destroy(b)
    params = twophase_params(nx, ny, nz)
    if nz == 1:   # a 2D grid for the oracle
        params = params.replace("grid_dim=3", "grid_dim=2")
    params = params.encode()
    Fo, rpo, cio, vo = oracle_residual(oracle_api, params, perm, poro, sc, sv, stage, p, s0, sw, 0.8)
    nnz = info[1]
    assert nnz == len(cio) and np.array_equal(rp, rpo) and np.array_equal(ci[:nnz], cio)
    assert np.allclose(F, Fo, rtol=1e-10, atol=1e-12 * np.abs(Fo).max())
    assert np.allclose(v[:nnz], vo, rtol=1e-10, atol=1e-12 * np.abs(vo).max())


# BiCGStab on the singular (no-flow) pressure system stagnates well above rounding, so
# the solves run to 1e-10 (FP64, and FP32 with Jacobi) and the results are compared at
# 1e-7; FP32 with ILU(0) runs at the benchmark's tolerances (Newton 1e-8, linear 1e-6),
# compared at 1e-5.
@pytest.mark.parametrize("params,newton_tol,solver_tol,bar", [
    # pressure by CG with the multigrid V-cycle (the default: full form)
    ("precision=fp64 preconditioner=ilu0", 1e-10, 1e-10, 1e-7),
    # pressure by CG with the Schur complement's diagonal (reduced form)
    ("precision=fp64 preconditioner=ilu0 krylov_form_twophase=reduced", 1e-10, 1e-10, 1e-7),
    # pressure by BiCGStab: reduced form, full form with red-black ILU(0)
    ("precision=fp64 preconditioner=ilu0 twophase_pressure_cg=0", 1e-10, 1e-10, 1e-7),
    ("precision=fp64 preconditioner=ilu0 twophase_pressure_mg=0 krylov_form_twophase=full", 1e-10, 1e-10, 1e-7),
    ("precision=fp64 preconditioner=diagonal", 1e-10, 1e-10, 1e-7),
    ("precision=mixed preconditioner=diagonal", 1e-10, 1e-10, 1e-7),
    ("precision=mixed preconditioner=ilu0", 1e-8, 1e-6, 1e-5),
    ("precision=mixed preconditioner=ilu0 twophase_pressure_mg=0", 1e-8, 1e-6, 1e-5)])
def test_fused_twophase_steps_match_oracle(hip_api, oracle_api, params, newton_tol, solver_tol, bar):
    nx, ny, nz = 12, 10, 6
    perm, poro, sw0, sc, sv = case(nx, ny, nz)
    n = perm.size
    dts = [1.0, 1.0, 2.0]
    po, so, _ = oracle_run(oracle_api, twophase_params(nx, ny, nz, rel_res_tol=1e-13, max_iter=40).encode(),
                           perm, poro, sw0, sc, sv, dts)
    b = fused_batch(hip_api, params + f" rel_res_tol={newton_tol} abs_res_tol=0 solver_tol={solver_tol} "
                                      "max_iter=40 solver_max_iter=5000")
    try:
        slot = fused_add(hip_api, b, nx, ny, nz, perm, poro, sw0, sc, sv)
        st = np.zeros(5, dtype=np.int32)
        tps = np.zeros(3, dtype=np.int32)
        for dt in dts:
            hip_api.batch_step(b, iptr(np.array([slot], dtype=np.int32)), 1, dt)
            hip_api.batch_stats(b, slot, iptr(st), None)
            hip_api.batch_twophase_stats(b, slot, iptr(tps))
            assert st[3] in (1, 2) and tps[2] in (1, 2) and st[4] == 0, (st, tps)
        ph, sh = np.zeros(n), np.zeros(n)
        hip_api.batch_twophase_state(b, slot, dptr(ph), dptr(sh))
    finally:
        hip_api.batch_destroy(b)
    assert np.allclose(sh, so, rtol=bar, atol=0.1 * bar)
    assert np.allclose(ph - ph[-1], po - po[-1], rtol=bar, atol=bar * np.abs(po).max())


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_multigrid_levels_above_the_tail_match_bicgstab(hip_api, precision):
    """24x20x16 (7680 cells): the fine level runs the one-level multigrid kernels, the
    levels from 12x10x8 down the workgroup tail.  CG with the V-cycle and BiCGStab with
    the red-black ILU(0) reach the same steps."""
    nx, ny, nz = 24, 20, 16
    perm, poro, sw0, sc, sv = case(nx, ny, nz, seed=3)
    n = perm.size
    out = {}
    for mg in (1, 0):
        b = fused_batch(hip_api, f"precision={precision} preconditioner=ilu0 rel_res_tol=1e-10 abs_res_tol=0 "
                                 f"solver_tol=1e-10 max_iter=40 solver_max_iter=5000 twophase_pressure_mg={mg} "
                                 "krylov_form_twophase=full" + ("" if mg else " twophase_pressure_cg=0"))
        try:
            slot = fused_add(hip_api, b, nx, ny, nz, perm, poro, sw0, sc, sv)
            st, tps = np.zeros(5, dtype=np.int32), np.zeros(3, dtype=np.int32)
            for dt in (1.0, 1.0):
                hip_api.batch_step(b, iptr(np.array([slot], dtype=np.int32)), 1, dt)
                hip_api.batch_stats(b, slot, iptr(st), None)
                hip_api.batch_twophase_stats(b, slot, iptr(tps))
                assert st[3] in (1, 2) and tps[2] in (1, 2) and st[4] == 0, (mg, st, tps)
            ph, sh = np.zeros(n), np.zeros(n)
            hip_api.batch_twophase_state(b, slot, dptr(ph), dptr(sh))
            out[mg] = (ph - ph.mean(), sh, int(tps[1]))
        finally:
            hip_api.batch_destroy(b)
    (p1, s1, it1), (p0, s0, it0) = out[1], out[0]
    assert np.allclose(s1, s0, rtol=1e-7, atol=1e-8)
    assert np.allclose(p1, p0, rtol=1e-7, atol=1e-7 * np.abs(p0).max())
    assert it1 < it0, (it1, it0)   # fewer pressure iterations with the V-cycle


@pytest.mark.parametrize("precision,newton_tol,solver_tol,bar", [("fp64", 1e-10, 1e-10, 1e-7),
                                                                  ("mixed", 1e-8, 1e-6, 1e-5)])
def test_multigrid_levels_above_the_tail_match_oracle(hip_api, oracle_api, precision, newton_tol, solver_tol, bar):
    """18x16x16 (4608 cells, above the 4096-cell workgroup tail): the fine level of the
    pressure V-cycle runs the one-level multigrid kernels (smoothing, restriction and
    prolongation folded into the up-sweep), the 9x8x8 level and below the tail.  Two steps
    of the default pressure stage (CG with the V-cycle) against the oracle's BiCGStab."""
    nx, ny, nz = 18, 16, 16
    perm, poro, sw0, sc, sv = case(nx, ny, nz, seed=3)
    n = perm.size
    dts = [1.0, 1.0]
    po, so, _ = oracle_run(oracle_api, twophase_params(nx, ny, nz, rel_res_tol=1e-11, max_iter=40).encode(),
                           perm, poro, sw0, sc, sv, dts)
    b = fused_batch(hip_api, f"precision={precision} preconditioner=ilu0 rel_res_tol={newton_tol} abs_res_tol=0 "
                             f"solver_tol={solver_tol} max_iter=40 solver_max_iter=5000 twophase_pressure_mg=1 "
                             "krylov_form_twophase=full")
    try:
        slot = fused_add(hip_api, b, nx, ny, nz, perm, poro, sw0, sc, sv)
        st, tps = np.zeros(5, dtype=np.int32), np.zeros(3, dtype=np.int32)
        for dt in dts:
            hip_api.batch_step(b, iptr(np.array([slot], dtype=np.int32)), 1, dt)
            hip_api.batch_stats(b, slot, iptr(st), None)
            hip_api.batch_twophase_stats(b, slot, iptr(tps))
            assert st[3] in (1, 2) and tps[2] in (1, 2) and st[4] == 0, (st, tps)
        ph, sh = np.zeros(n), np.zeros(n)
        hip_api.batch_twophase_state(b, slot, dptr(ph), dptr(sh))
    finally:
        hip_api.batch_destroy(b)
    assert np.allclose(sh, so, rtol=bar, atol=0.1 * bar)
    assert np.allclose(ph - ph[-1], po - po[-1], rtol=bar, atol=bar * np.abs(po).max())


@pytest.mark.parametrize("precision", ["mixed", "fp64"])   # fp64: both kinds in the reduced form
def test_fused_twophase_and_heat_share_a_batch(hip_api, precision):
    """Two-phase and heat instances in one batch give bit-identical results to each run
    alone (per-instance reductions, stage-independent launches)."""
    from equelle_amd.workloads import gaussian_bumps
    params = f"precision={precision} preconditioner=ilu0 rel_res_tol=1e-8 solver_tol=1e-6 max_iter=30"
    tp = [(10, 9, 7), (20, 15, 4)]
    heat = [(64, 40, 0.2)]

    def run(which):
        b = fused_batch(hip_api, params)
        try:
            slots = []
            for q, (nx, ny, nz) in enumerate(tp):
                if which in ("all", f"tp{q}"):
                    perm, poro, sw0, sc, sv = case(nx, ny, nz, seed=q)
                    slots.append(("tp", nx * ny * nz, fused_add(hip_api, b, nx, ny, nz, perm, poro, sw0, sc, sv)))
            for q, (nx, ny, k) in enumerate(heat):
                if which in ("all", f"heat{q}"):
                    s = C.c_int()
                    hip_api.batch_add(b, nx, ny, k, 1.0, dptr(gaussian_bumps(nx, ny, 4, width_scale=0.05)),
                                      C.byref(s))
                    slots.append(("heat", nx * ny, s.value))
            arr = np.array([s for _, _, s in slots], dtype=np.int32)
            for dt in (1.0, 0.5):
                hip_api.batch_step(b, iptr(arr), len(arr), dt)
            out = []
            for kind, n, s in slots:
                u = np.zeros(n)
                hip_api.batch_state(b, s, dptr(u))
                out.append(u)
            return out
        finally:
            hip_api.batch_destroy(b)

    together = run("all")
    alone = run("tp0") + run("tp1") + run("heat0")
    for a, c in zip(together, alone):
        assert np.array_equal(a, c)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_fused_twophase_speculative_matches_checked(hip_api, precision):
    """Two-phase instances (multigrid CG pressure stage, reduced-form transport stage) give
    the same steps whether the work after each Krylov solve waits for the convergence count
    or is queued behind the predicted iterations."""
    cases = [(24, 20, 16, 3), (16, 12, 8, 4)]
    data = [case(nx, ny, nz, seed=sd) for nx, ny, nz, sd in cases]
    res = {}
    for spec in (0, 1):
        b = fused_batch(hip_api, f"precision={precision} preconditioner=ilu0 rel_res_tol=1e-9 abs_res_tol=0 "
                                 f"solver_tol=1e-8 max_iter=40 solver_max_iter=5000 graph=1 speculate={spec}")
        try:
            slots = np.array([fused_add(hip_api, b, nx, ny, nz, *d) for (nx, ny, nz, _), d in zip(cases, data)],
                             dtype=np.int32)
            st, tps = np.zeros(5, dtype=np.int32), np.zeros(3, dtype=np.int32)
            traj = []
            for dt in (1.0, 2.0, 2.0):
                hip_api.batch_step(b, iptr(slots), len(slots), dt)
                for sl, (nx, ny, nz, _) in zip(slots, cases):
                    hip_api.batch_stats(b, int(sl), iptr(st), None)
                    hip_api.batch_twophase_stats(b, int(sl), iptr(tps))
                    assert st[3] in (1, 2) and tps[2] in (1, 2) and st[4] == 0, (spec, st, tps)
                    ph, sh = np.zeros(nx * ny * nz), np.zeros(nx * ny * nz)
                    hip_api.batch_twophase_state(b, int(sl), dptr(ph), dptr(sh))
                    traj.append((ph, sh, st[:3].tolist(), tps.tolist()))
            res[spec] = traj
        finally:
            hip_api.batch_destroy(b)
    for a, g in zip(res[0], res[1]):
        assert a[2] == g[2] and a[3] == g[3]
        assert np.array_equal(a[0], g[0]) and np.array_equal(a[1], g[1])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_fused_twophase_split_form_matches_full_form(hip_api, precision):
    """The transport stage in the reduced form inside the full form's storage (the
    default with the multigrid pressure stage) against the full form with red-black
    ILU(0): equal in exact arithmetic, so the steps agree to the solver tolerances."""
    nx, ny, nz = 24, 20, 16
    perm, poro, sw0, sc, sv = case(nx, ny, nz, seed=7)
    n = perm.size
    out = {}
    for split in (0, 1):
        b = fused_batch(hip_api, f"precision={precision} preconditioner=ilu0 rel_res_tol=1e-10 abs_res_tol=0 "
                                 f"solver_tol=1e-10 max_iter=40 solver_max_iter=5000 twophase_split_form={split}")
        try:
            slot = fused_add(hip_api, b, nx, ny, nz, perm, poro, sw0, sc, sv)
            st, tps = np.zeros(5, dtype=np.int32), np.zeros(3, dtype=np.int32)
            for dt in (1.0, 1.0):
                hip_api.batch_step(b, iptr(np.array([slot], dtype=np.int32)), 1, dt)
                hip_api.batch_stats(b, slot, iptr(st), None)
                hip_api.batch_twophase_stats(b, slot, iptr(tps))
                assert st[3] in (1, 2) and tps[2] in (1, 2) and st[4] == 0, (split, st, tps)
            ph, sh = np.zeros(n), np.zeros(n)
            hip_api.batch_twophase_state(b, slot, dptr(ph), dptr(sh))
            out[split] = (ph - ph.mean(), sh)
        finally:
            hip_api.batch_destroy(b)
    (p1, s1), (p0, s0) = out[1], out[0]
    assert np.allclose(s1, s0, rtol=1e-7, atol=1e-8)
    assert np.allclose(p1, p0, rtol=1e-7, atol=1e-7 * np.abs(p0).max())


def test_fully_implicit_cpr_needs_fewer_gmres_iterations(hip_api):
    """CPR (one multigrid V-cycle on the pressure block, then ILU(0) of the coupled system)
    against ILU(0) alone: the same Newton solution (1e-9) with fewer GMRES iterations on a
    grid whose pressure block has multigrid levels above the one-workgroup tail."""
    nx, ny, nz = 24, 20, 16
    perm, poro, sw0, sc, sv = case(nx, ny, nz)
    n = perm.size
    out = {}
    for pc in ("ILU0", "cpr"):
        params = fi_params(nx, ny, nz, rel_res_tol=1e-12, max_iter=40, solver_tol=1e-7,
                           extra=f"precision=fp64 solver=GMRes preconditioner={pc} fully_implicit=1").encode()
        sim = create(hip_api, params, perm, poro, sw0, sc, sv)
        try:
            st, gm = np.zeros(6, dtype=np.int32), 0
            for dt in (0.05, 0.05):
                hip_api.twophase_step(sim, dt, iptr(st), None)
                assert st[2] in (1, 2), st
                gm += int(st[1])
            ph, sh = np.zeros(n), np.zeros(n)
            hip_api.twophase_state(sim, dptr(ph), dptr(sh))
        finally:
            hip_api.twophase_destroy(sim)
        out[pc] = (ph, sh, gm)
    assert np.allclose(out["cpr"][1], out["ILU0"][1], rtol=1e-9, atol=1e-10)
    assert out["cpr"][2] < out["ILU0"][2], (out["cpr"][2], out["ILU0"][2])
