"""Pins the oracle's two-phase program (examples/dsl/twophase.equelle, Corey variant)
by mathematics a plausible mistake would break:

* transmissibility: for unit cells the program's half-transmissibility harmonic mean
  is the closed form 2 k1 k2 / (k1 + k2), positive (a sign error in either half
  transmissibility breaks it); checked through the pressure Jacobian's off-diagonals;
* pressure Jacobian: symmetric, rows sum to zero (no-flow boundary, constants in the
  null space), and the first step is linear (one Newton iteration);
* transport Jacobian: equals central finite differences of the residual (every AD
  rule in the residual: products, quotient, upwind trinaryIf, On, divergence);
* mass balance: sum pv (sw - sw0) = dt * sum(q) at convergence, q = injection +
  outflow * fractional flow (divergence telescopes);
* the pressure drives water from injector to producer (p_inj > p_prod) and the
  saturation stays in [0, 1].
"""
import numpy as np

from equelle_cases import csr_dense
from twophase_cases import case, oracle_residual, oracle_run, twophase_params

NX, NY, NZ = 7, 6, 4


def test_pressure_jacobian_closed_form(oracle_api):
    perm, poro, sw0, sc, sv = case(NX, NY, NZ)
    n = perm.size
    params = twophase_params(NX, NY, NZ).encode()
    p = np.random.default_rng(1).uniform(0, 1, n)
    F, rp, ci, v = oracle_residual(oracle_api, params, perm, poro, sc, sv, 0, p, sw0, sw0, 1.0)
    J = csr_dense(rp, ci, v, n, n)
    assert np.allclose(J, J.T, rtol=1e-13, atol=0)
    assert np.allclose(J.sum(axis=1), 0.0, atol=1e-10 * np.abs(J).max())
    # sw0 = 0: total mobility = 1/oilvisc everywhere; J[c, c+1] = -trans * lambda
    lam = 1.0 / 5e-3
    for c in range(n - 1):
        if (c % NX) != NX - 1:
            t = 2 * perm[c] * perm[c + 1] / (perm[c] + perm[c + 1])
            assert J[c, c + 1] < 0 and abs(-J[c, c + 1] - t * lam) <= 1e-12 * t * lam
    # F(p) = J p - source (linear in p for a fixed upwind direction)
    src = np.zeros(n)
    src[sc] = sv
    assert np.allclose(F, J @ p - src, rtol=1e-10, atol=1e-10 * np.abs(F).max())


def test_transport_jacobian_matches_finite_differences(oracle_api):
    perm, poro, sw0, sc, sv = case(NX, NY, NZ)
    n = perm.size
    params = twophase_params(NX, NY, NZ).encode()
    rng = np.random.default_rng(3)
    p = rng.uniform(0, 1, n)
    s0 = rng.uniform(0.05, 0.9, n)
    sw = rng.uniform(0.1, 0.9, n)
    F, rp, ci, v = oracle_residual(oracle_api, params, perm, poro, sc, sv, 1, p, s0, sw, 0.7)
    J = csr_dense(rp, ci, v, n, n)
    h = 1e-6
    for col in rng.choice(n, 25, replace=False):
        e = np.zeros(n)
        e[col] = h
        Fp = oracle_residual(oracle_api, params, perm, poro, sc, sv, 1, p, s0, sw + e, 0.7)[0]
        Fm = oracle_residual(oracle_api, params, perm, poro, sc, sv, 1, p, s0, sw - e, 0.7)[0]
        fd = (Fp - Fm) / (2 * h)
        assert np.allclose(J[:, col], fd, rtol=1e-6, atol=1e-7 * max(1.0, np.abs(fd).max())), col


def test_steps_conserve_mass_and_stay_physical(oracle_api):
    perm, poro, sw0, sc, sv = case(NX, NY, NZ)
    params = twophase_params(NX, NY, NZ, rel_res_tol=1e-13, max_iter=40).encode()
    p1, s1, it1 = oracle_run(oracle_api, params, perm, poro, sw0, sc, sv, [1.0])
    assert it1[0] == 1                          # first pressure solve: linear, one Newton step
    p2, s2, _ = oracle_run(oracle_api, params, perm, poro, sw0, sc, sv, [1.0, 1.0])
    pv = 0.2
    # step 2: q = injection + outflow * f(sw), f = mw / (mw + mo) at the new saturation
    mw = s2 ** 2 / 1e-3
    mo = (1 - s2) ** 2 / 5e-3
    f = mw / (mw + mo)
    q = np.zeros_like(s2)
    inj, prod = sc[sv > 0], sc[sv < 0]
    q[inj] = sv[sv > 0]
    q[prod] = sv[sv < 0] * f[prod]
    assert abs(pv * (s2 - s1).sum() - 1.0 * q.sum()) <= 1e-10 * np.abs(q).sum()
    assert s2.min() >= -1e-12 and s2.max() <= 1 + 1e-12
    assert p2[inj].mean() > p2[prod].mean()
