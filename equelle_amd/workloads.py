"""Seeded synthetic inputs for the benchmark and test workloads.

Every unspecified choice is frozen here (and only here):

linear heat (100 x 100)      u0 ~ U[0,1] (seed 0); Dirichlet u=0 on the x=0 faces
                              (ids j*(nx+1)), u=1 on the x=max faces (ids nx + j*(nx+1));
                              k = 0.3, cv = 1, dt = 0.5, 10 steps.
nonlinear heat (2048^2)      k(u) = 0.1 (1 + u^2), face value = mean of the two cells;
                              u0 = sum of 16 Gaussian bumps A exp(-r^2 / (2 w^2)) with
                              A ~ U[0.5, 2], w ~ U[5, 50] cells, centres ~ U(domain), plus
                              U[0, 0.01] noise; zero-flux boundary; dt = 1.0, 20 steps;
                              Newton to ||F(u_k)|| <= 1e-8 ||F(u_0)||.
two-phase (128 x 128 x 64)   log k ~ N(0,1) filtered by a Gaussian of sigma 4/3 cells truncated at
                              radius 4 (scipy.ndimage, reflecting edges), renormalised to mean 0 and
                              variance 1; porosity 0.2; sw_initial = 0; injector column (0,0,k) and
                              producer column (nx-1,ny-1,k), +-1 per cell; Corey exponent 2;
                              watervisc 1e-3, oilvisc 5e-3 (ratio 5); dt = 1.0, 30 steps; each
                              Newton solve to 1e-8 relative residual.
batched heat (512 instances) side ~ round(exp(U[ln 32, ln 512])), k0 ~ U[0.05, 0.5],
                              bumps scaled to the instance (width U[5,50] * side/2048, at
                              least 1.5 cells), seed = base + instance.
"""
from __future__ import annotations

import numpy as np

TWOPHASE = dict(nx=128, ny=128, nz=64, dt=1.0, steps=30, watervisc=1e-3, oilvisc=5e-3, corey=2, rate=1.0,
                porosity=0.2, rel_res_tol=1e-8)
NONLINEAR_HEAT = dict(nx=2048, ny=2048, k=0.1, cv=1.0, dt=1.0, steps=20, rel_res_tol=1e-8)
LINEAR_HEAT = dict(nx=100, ny=100, k=0.3, cv=1.0, dt=0.5, steps=10)


def gaussian_bumps(nx: int, ny: int, seed: int, nbumps: int = 16, width_scale: float = 1.0) -> np.ndarray:
    """u0 on cell centres (x-fastest cell order, unit cells)."""
    rng = np.random.default_rng(seed)
    amp = rng.uniform(0.5, 2.0, nbumps)
    width = np.maximum(rng.uniform(5.0, 50.0, nbumps) * width_scale, 1.5)
    cx = rng.uniform(0.0, nx, nbumps)
    cy = rng.uniform(0.0, ny, nbumps)
    x = np.arange(nx) + 0.5
    y = np.arange(ny) + 0.5
    u = np.zeros((ny, nx))
    for a, w, px, py in zip(amp, width, cx, cy):
        gx = np.exp(-((x - px) ** 2) / (2 * w * w))
        gy = np.exp(-((y - py) ** 2) / (2 * w * w))
        u += a * np.outer(gy, gx)
    u += rng.uniform(0.0, 0.01, (ny, nx))
    return np.ascontiguousarray(u.ravel())


def linear_heat_inputs(nx: int = 100, ny: int = 100, seed: int = 0):
    u0 = np.random.default_rng(seed).uniform(0.0, 1.0, nx * ny)
    faces, vals = [], []
    for j in range(ny):
        faces += [j * (nx + 1), nx + j * (nx + 1)]
        vals += [0.0, 1.0]
    return u0, np.array(faces, dtype=np.int32), np.array(vals)


def batched_instances(count: int = 512, seed: int = 1234, lo: int = 32, hi: int = 512):
    """[(side, k0, u0)] for the batched nonlinear-heat workload."""
    rng = np.random.default_rng(seed)
    sides = np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), count))).astype(int)
    k0 = rng.uniform(0.05, 0.5, count)
    out = []
    for i, (s, k) in enumerate(zip(sides, k0)):
        out.append((int(s), float(k), gaussian_bumps(int(s), int(s), seed + 1 + i, width_scale=s / 2048.0)))
    return out


def twophase_inputs(nx: int, ny: int, nz: int, seed: int = 42, rate: float = 1.0, porosity: float = 0.2):
    """perm, poro, sw_initial (cell order x fastest), source cells (ascending), source rates."""
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((nz, ny, nx))
    f = gaussian_filter(g, sigma=4.0 / 3.0, truncate=3.0, mode="reflect")
    f = (f - f.mean()) / f.std()
    perm = np.ascontiguousarray(np.exp(f).ravel())
    n = nx * ny * nz
    poro = np.full(n, porosity)
    sw0 = np.zeros(n)
    cells, vals = [], []
    for k in range(nz):
        cells += [k * nx * ny, k * nx * ny + nx * ny - 1]
        vals += [rate, -rate]
    if nx * ny == 1:
        raise ValueError("two-phase grid needs at least two columns")
    return perm, poro, sw0, np.array(cells, dtype=np.int32), np.array(vals)


def twophase_params(nx, ny, nz, precision="fp64", solver="GMRes", precond="diagonal", rel_res_tol=1e-8,
                    solver_tol=1e-6, max_iter=20, extra="") -> str:
    t = TWOPHASE
    return (f"grid_dim=3 nx={nx} ny={ny} nz={nz} dx=1 dy=1 dz=1 watervisc={t['watervisc']} oilvisc={t['oilvisc']} "
            f"corey={t['corey']} abs_res_tol=0 rel_res_tol={rel_res_tol} max_iter={max_iter} solver={solver} "
            f"preconditioner={precond} precision={precision} solver_tol={solver_tol} solver_max_iter=5000 "
            f"gmres_restart=30 {extra}").strip()


def heat_params(nx, ny, k, cv=1.0, quadratic=True, precision="fp64", precond="ilu0", solver="BiCGStab",
                rel_res_tol=1e-8, solver_tol=1e-6, max_iter=20, extra="") -> str:
    return (f"grid_dim=2 nx={nx} ny={ny} dx=1 dy=1 k={k} cv={cv} "
            f"k_model={'quadratic' if quadratic else 'const'} abs_res_tol=0 rel_res_tol={rel_res_tol} "
            f"max_iter={max_iter} solver={solver} preconditioner={precond} precision={precision} "
            f"solver_tol={solver_tol} solver_max_iter=2000 {extra}").strip()
