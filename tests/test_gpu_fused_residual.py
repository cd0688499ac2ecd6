"""GPU: recorded residual programs against the op-by-op AD engine and the oracle.

newtonSolve records the operations of its first residual evaluation and runs every later
evaluation as one fused kernel (fuse_residual=1, the default); fuse_residual=0 keeps the
op-by-op evaluation of the reference (EquelleRuntimeCUDA_impl.hpp:216-300).  Bars:
* each kernel is checked once per grid against the recorded evaluation (values and
  Jacobian, 1e-10 relative) before it replaces it; no kernel may be rejected;
* Newton trajectories of the fused and the op-by-op runs agree to 1e-10 with the same
  iteration counts, and both meet the oracle at the bars of the other parity tests;
* later newtonSolve calls at the same call site record symbolically and find their kernel.
"""
import ctypes as C

import numpy as np
import pytest

from equelle_amd.capi import dptr, iptr
from equelle_cases import c0_dirichlet
from twophase_cases import case, fi_params, oracle_fi_run, oracle_run, twophase_params

pytestmark = pytest.mark.gpu


def stats(api):
    out, secs = (C.c_long * 6)(), C.c_double()
    assert api.residual_stats(out, C.byref(secs)) == 0
    return dict(zip(("recorded", "symbolic", "compiled", "evaluations", "aborted", "rejected"), list(out)))


def close(a, b, rtol):
    a, b = np.asarray(a), np.asarray(b)
    return np.allclose(a, b, rtol=rtol, atol=rtol * max(1.0, float(np.abs(b).max())))


def heat_run(api, params, u0, dirf, dirv, dts):
    n = u0.size
    u, it = np.zeros(n), np.zeros(len(dts), dtype=np.int32)
    api.heat_run(params.encode(), dptr(u0), n, iptr(dirf), dptr(dirv), len(dirf), dptr(dts), len(dts), dptr(u),
                 iptr(it), None)
    return u, it


@pytest.mark.parametrize("grid", ["grid_dim=2 nx=96 ny=80", "grid_dim=2 nx=61 ny=47", "grid_dim=3 nx=20 ny=14 nz=9"])
@pytest.mark.parametrize("k_model", ["quadratic", "const"])
def test_fused_heat_matches_op_by_op_and_oracle(hip_api, oracle_api, grid, k_model):
    dims = dict(kv.split("=") for kv in grid.split())
    nx, ny, nz = int(dims["nx"]), int(dims["ny"]), int(dims.get("nz", 1))
    n = nx * ny * nz
    if nz == 1:
        dirf, dirv = c0_dirichlet(nx, ny)
    else:
        dirf, dirv = np.zeros(1, dtype=np.int32)[:0], np.zeros(0)
    base = (f"{grid} k=0.1 cv=1 k_model={k_model} abs_res_tol=0 rel_res_tol=1e-10 max_iter=20 "
            "solver=BiCGStab preconditioner=ilu0 solver_tol=1e-11 solver_max_iter=3000")
    rng = np.random.default_rng(n)
    u0 = rng.uniform(0, 2, n)
    dts = np.array([1.0, 0.5, 1.0])
    s0 = stats(hip_api)
    uf, itf = heat_run(hip_api, base, u0, dirf, dirv, dts)
    s1 = stats(hip_api)
    ue, ite = heat_run(hip_api, base + " fuse_residual=0", u0, dirf, dirv, dts)
    s2 = stats(hip_api)
    uo, _ = heat_run(oracle_api, base, u0, dirf, dirv, dts)
    assert s1["rejected"] == s0["rejected"] and s1["aborted"] == s0["aborted"], (s0, s1)
    assert s1["evaluations"] > s0["evaluations"], (s0, s1)
    # steps 2 and 3 find the kernel through a symbolic recording
    assert s1["symbolic"] >= s0["symbolic"] + 2, (s0, s1)
    assert s2["evaluations"] == s1["evaluations"], "fuse_residual=0 must evaluate op by op"
    assert np.array_equal(itf, ite), (itf, ite)
    assert close(uf, ue, 1e-10)
    assert close(uf, uo, 1e-8)


@pytest.mark.parametrize("extra", ["precision=fp64 solver=GMRes preconditioner=diagonal",
                                   "precision=mixed solver=GMRes preconditioner=diagonal"])
def test_fused_twophase_program_matches_op_by_op_and_oracle(hip_api, oracle_api, extra):
    """The two-phase program's pressure and transport residuals (upwind trinaryIf on the
    gradient's sign, Corey mobilities, sources, On / Extend of the well cells)."""
    nx, ny, nz = 12, 10, 6
    perm, poro, sw0, sc, sv = case(nx, ny, nz)
    n = perm.size
    dts = [1.0, 1.0, 2.0]
    out = {}
    s0 = stats(hip_api)
    for fuse in (1, 0):
        params = twophase_params(nx, ny, nz, rel_res_tol=1e-12, max_iter=40, solver_tol=1e-13,
                                 extra=extra + f" fuse_residual={fuse}").encode()
        sim = C.c_void_p()
        hip_api.twophase_create(params, dptr(perm), dptr(poro), dptr(sw0), n, iptr(sc), dptr(sv), len(sc),
                                C.byref(sim))
        try:
            st, its = np.zeros(6, dtype=np.int32), []
            for dt in dts:
                hip_api.twophase_step(sim, dt, iptr(st), None)
                its.append((st[0], st[3]))
            ph, sh = np.zeros(n), np.zeros(n)
            hip_api.twophase_state(sim, dptr(ph), dptr(sh))
        finally:
            hip_api.twophase_destroy(sim)
        out[fuse] = (ph, sh, its)
    s1 = stats(hip_api)
    assert s1["rejected"] == s0["rejected"] and s1["aborted"] == s0["aborted"], (s0, s1)
    assert s1["evaluations"] > s0["evaluations"], (s0, s1)
    (pf, sf, itf), (pe, se, ite) = out[1], out[0]
    assert itf == ite, (itf, ite)
    assert close(sf, se, 1e-10)
    assert close(pf - pf[-1], pe - pe[-1], 1e-9)
    oparams = twophase_params(nx, ny, nz, rel_res_tol=1e-13, max_iter=40).encode()
    po, so, _ = oracle_run(oracle_api, oparams, perm, poro, sw0, sc, sv, dts)
    assert close(sf, so, 1e-9)


def test_fused_fully_implicit_matches_op_by_op_and_oracle(hip_api, oracle_api):
    """newtonSolveSystem: the stacked unknown's subset / superset ranges (On / Extend of
    range sets) and the coupled 2n x 2n Jacobian rows in one kernel."""
    nx, ny, nz = 8, 7, 4
    perm, poro, sw0, sc, sv = case(nx, ny, nz)
    n = perm.size
    dts = [0.05, 0.05, 0.1]
    out = {}
    s0 = stats(hip_api)
    for fuse in (1, 0):
        params = fi_params(nx, ny, nz, rel_res_tol=1e-12, max_iter=40, solver_tol=1e-7,
                           extra=f"precision=fp64 solver=GMRes preconditioner=ILU0 fully_implicit=1 "
                                 f"fuse_residual={fuse}").encode()
        sim = C.c_void_p()
        hip_api.twophase_create(params, dptr(perm), dptr(poro), dptr(sw0), n, iptr(sc), dptr(sv), len(sc),
                                C.byref(sim))
        try:
            st = np.zeros(6, dtype=np.int32)
            for dt in dts:
                hip_api.twophase_step(sim, dt, iptr(st), None)
            ph, sh = np.zeros(n), np.zeros(n)
            hip_api.twophase_state(sim, dptr(ph), dptr(sh))
        finally:
            hip_api.twophase_destroy(sim)
        out[fuse] = (ph, sh)
    s1 = stats(hip_api)
    assert s1["rejected"] == s0["rejected"] and s1["aborted"] == s0["aborted"], (s0, s1)
    assert s1["evaluations"] > s0["evaluations"], (s0, s1)
    assert close(out[1][1], out[0][1], 1e-9)
    oparams = fi_params(nx, ny, nz, rel_res_tol=1e-13, max_iter=40).encode()
    po, so, _ = oracle_fi_run(oracle_api, oparams, perm, poro, sw0, sc, sv, dts)
    assert close(out[1][1], so, 1e-9)
