"""Pins the oracle's fully implicit two-phase program
(examples/dsl/twophase_fully_implicit.equelle, newtonSolveSystem of
EquelleRuntimeCPU_impl.hpp:282-339) by mathematics a plausible mistake would break:

* the combined 2n x 2n Jacobian equals central finite differences of the combined
  residual in all four blocks (pressure and transport rows, p and sw columns): pins
  subset/superset of the combined primary variable, the mobility coupling dF_p/dsw and
  the flux coupling dF_s/dp;
* null space: a constant pressure shift changes nothing (J [1; 0] = 0) and the pressure
  rows sum to zero for any sw (1^T J_p. = 0: divergence telescopes);
* where the two programs coincide (sw = sw0) the residual equals the sequential
  program's pressure and transport residuals;
* a converged step conserves water: sum pv (sw - sw0) = dt sum q.
"""
import numpy as np

from equelle_cases import csr_dense
from twophase_cases import case, fi_params, oracle_fi_residual, oracle_fi_run, oracle_residual

NX, NY, NZ = 5, 4, 3


def _fi(oracle_api, p, sw0, sw, dt=0.6):
    perm, poro, _, sc, sv = case(NX, NY, NZ)
    params = fi_params(NX, NY, NZ).encode()
    return oracle_fi_residual(oracle_api, params, perm, poro, sc, sv, p, sw0, sw, dt)


def test_fi_jacobian_matches_finite_differences(oracle_api):
    n = NX * NY * NZ
    rng = np.random.default_rng(11)
    p, s0, sw = rng.uniform(0, 1, n), rng.uniform(0.05, 0.9, n), rng.uniform(0.1, 0.9, n)
    F, rp, ci, v = _fi(oracle_api, p, s0, sw)
    assert F.size == 2 * n
    J = csr_dense(rp, ci, v, 2 * n, 2 * n)
    h = 1e-6
    for col in list(rng.choice(n, 12, replace=False)) + list(n + rng.choice(n, 12, replace=False)):
        e = np.zeros(2 * n)
        e[col] = h
        Fp = _fi(oracle_api, p + e[:n], s0, sw + e[n:])[0]
        Fm = _fi(oracle_api, p - e[:n], s0, sw - e[n:])[0]
        fd = (Fp - Fm) / (2 * h)
        # upwind switches are not differentiable; the random state keeps all faces away
        # from p1 == p2, so the central difference sees one branch
        assert np.allclose(J[:, col], fd, rtol=1e-5, atol=1e-6 * np.abs(J).max()), col
    # both coupling blocks are present
    assert np.abs(J[:n, n:]).max() > 0 and np.abs(J[n:, :n]).max() > 0


def test_fi_null_space(oracle_api):
    n = NX * NY * NZ
    rng = np.random.default_rng(12)
    p, s0, sw = rng.uniform(0, 1, n), rng.uniform(0.05, 0.9, n), rng.uniform(0.1, 0.9, n)
    F, rp, ci, v = _fi(oracle_api, p, s0, sw)
    J = csr_dense(rp, ci, v, 2 * n, 2 * n)
    one_p = np.concatenate([np.ones(n), np.zeros(n)])
    assert np.abs(J @ one_p).max() <= 1e-10 * np.abs(J).max()
    assert np.abs(one_p @ J).max() <= 1e-10 * np.abs(J).max()
    F2 = _fi(oracle_api, p + 3.0, s0, sw)[0]
    assert np.allclose(F2, F, rtol=1e-12, atol=1e-12 * np.abs(F).max())


def test_fi_residual_equals_sequential_where_they_coincide(oracle_api):
    perm, poro, _, sc, sv = case(NX, NY, NZ)
    params = fi_params(NX, NY, NZ).encode()
    n = perm.size
    rng = np.random.default_rng(13)
    p, s = rng.uniform(0, 1, n), rng.uniform(0.1, 0.9, n)
    F = oracle_fi_residual(oracle_api, params, perm, poro, sc, sv, p, s, s, 0.8)[0]
    Fp = oracle_residual(oracle_api, params, perm, poro, sc, sv, 0, p, s, s, 0.8)[0]
    Fs = oracle_residual(oracle_api, params, perm, poro, sc, sv, 1, p, s, s, 0.8)[0]
    assert np.allclose(F[:n], Fp, rtol=1e-13, atol=1e-13 * np.abs(Fp).max())
    assert np.allclose(F[n:], Fs, rtol=1e-13, atol=1e-13 * np.abs(Fs).max())


def test_fi_step_conserves_water(oracle_api):
    perm, poro, sw0, sc, sv = case(NX, NY, NZ)
    params = fi_params(NX, NY, NZ, rel_res_tol=1e-12, max_iter=40).encode()
    dt = 0.2
    p, sw, it = oracle_fi_run(oracle_api, params, perm, poro, sw0, sc, sv, [dt])
    assert it[0] >= 2                           # nonlinear in sw: more than one Newton step
    pv = 0.2                                    # unit cells, porosity 0.2
    mw, mo = sw / 1e-3, (1 - sw) / 5e-3         # Corey 1 (the program), the workload's viscosities
    f = mw / (mw + mo)
    q = np.zeros_like(sw)
    inj, prod = sc[sv > 0], sc[sv < 0]
    q[inj] = sv[sv > 0]
    q[prod] = sv[sv < 0] * f[prod]
    assert abs(pv * (sw - sw0).sum() - dt * q.sum()) <= 1e-9 * np.abs(q).sum()
    assert sw.min() >= -1e-12 and sw.max() <= 1 + 1e-12
    assert p[inj].mean() > p[prod].mean()
