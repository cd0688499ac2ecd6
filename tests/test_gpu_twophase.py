"""GPU parity of the two-phase program (examples/dsl/twophase.equelle with Corey
exponent 2) on the device AD engine against the oracle.

* pressure and transport residual + Jacobian: CSR pattern bit-exact, values 1e-10;
* time steps (pressure newtonSolve + saturation newtonSolve): saturation within
  1e-9 of the oracle's (GMRES; 1e-7 for BiCGStab), pressure within the same bar up to
  the additive constant the no-flow pressure equation leaves free (compared after
  subtracting the last cell's value);
* FP32 GMRES(30) + Jacobi with FP64 refinement converges to the FP64 result.
"""
import ctypes as C

import numpy as np
import pytest

from equelle_amd.capi import dptr, iptr
from twophase_cases import case, oracle_residual, oracle_run, twophase_params

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
