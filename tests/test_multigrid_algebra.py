"""The algebra of the fused engine's multigrid preconditioner for the two-phase pressure
stage, checked on the CPU against independently built dense matrices.

The device kernels (heat_fused.hip, k_mg_*) store every level in the red-black layout
whose index maps test_redblack_algebra.py pins; here the same operations are restated on
natural cell order with colour masks:

* aggregation: cell (i, j, k) of level l belongs to aggregate (i//2, j//2, k//2) of level
  l + 1 (ceil(n/2) cells per axis), P piecewise constant;
* the Galerkin coarse operator as the kernel forms it (per fine cell: the diagonal to the
  coarse diagonal, a coupling to the coarse diagonal when the neighbour is in the same
  aggregate, else to the coarse coupling in that direction) equals P^T A P, and is again a
  7-point stencil with zero row sums;
* the coarsest solve z = (A + s 1 1^T)^-1 b, s = mean diagonal / n, gives the zero-sum
  solution of A z = b for b with zero sum;
* one V-cycle (red then black Gauss-Seidel half sweeps, the red cells' residual summed
  per aggregate, the coarse correction, black then red half sweeps) is a symmetric
  positive definite operator, so CG applies, and CG with it needs several times fewer
  iterations than CG with the diagonal on a heterogeneous pressure matrix.
"""
import numpy as np
import pytest

COARSE = 64   # the largest level solved densely (MG_COARSE)


def idx(dims, i, j, k):
    nx, ny, _ = dims
    return i + nx * (j + ny * k)


def pressure_matrix(dims, rng, spread=3.0):
    """7-point symmetric M-matrix with zero row sums (no-flow boundary): a face of
    transmissibility t between cells a and b adds t to both diagonals and -t to both
    couplings; log-normal t."""
    nx, ny, nz = dims
    n = nx * ny * nz
    A = np.zeros((n, n))
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                c = idx(dims, i, j, k)
                for (di, dj, dk) in ((1, 0, 0), (0, 1, 0), (0, 0, 1)):
                    if i + di < nx and j + dj < ny and k + dk < nz:
                        m = idx(dims, i + di, j + dj, k + dk)
                        t = float(np.exp(spread * rng.standard_normal()))
                        A[c, c] += t
                        A[m, m] += t
                        A[c, m] -= t
                        A[m, c] -= t
    return A


def stencil(A, dims):
    """Coefficient arrays C, W, E, S, N, B, T per cell (0 where the neighbour is absent)."""
    nx, ny, nz = dims
    n = nx * ny * nz
    K = np.zeros((7, n))
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                c = idx(dims, i, j, k)
                K[0, c] = A[c, c]
                if i > 0:
                    K[1, c] = A[c, c - 1]
                if i < nx - 1:
                    K[2, c] = A[c, c + 1]
                if j > 0:
                    K[3, c] = A[c, c - nx]
                if j < ny - 1:
                    K[4, c] = A[c, c + nx]
                if k > 0:
                    K[5, c] = A[c, c - nx * ny]
                if k < nz - 1:
                    K[6, c] = A[c, c + nx * ny]
    return K


def dense(K, dims):
    nx, ny, nz = dims
    n = nx * ny * nz
    A = np.zeros((n, n))
    off = {1: -1, 2: 1, 3: -nx, 4: nx, 5: -nx * ny, 6: nx * ny}
    for c in range(n):
        A[c, c] += K[0, c]
        for q, o in off.items():
            if K[q, c] != 0.0:
                A[c, c + o] += K[q, c]
    return A


def coarse_dims(dims):
    return tuple((d + 1) // 2 for d in dims)


def prolongation(dims):
    cd = coarse_dims(dims)
    nx, ny, nz = dims
    P = np.zeros((nx * ny * nz, cd[0] * cd[1] * cd[2]))
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                P[idx(dims, i, j, k), idx(cd, i // 2, j // 2, k // 2)] = 1.0
    return P


def galerkin_rule(K, dims):
    """The coarse stencil as k_mg_galerkin forms it, one coarse cell at a time."""
    cd = coarse_dims(dims)
    nx, ny, nz = dims
    Kc = np.zeros((7, cd[0] * cd[1] * cd[2]))
    for kc in range(cd[2]):
        for jc in range(cd[1]):
            for ic in range(cd[0]):
                co = np.zeros(7)
                for k in range(2 * kc, min(2 * kc + 2, nz)):
                    for j in range(2 * jc, min(2 * jc + 2, ny)):
                        for i in range(2 * ic, min(2 * ic + 2, nx)):
                            s = idx(dims, i, j, k)
                            co[0] += K[0, s]
                            if i > 0:
                                co[0 if i & 1 else 1] += K[1, s]
                            if i < nx - 1:
                                co[2 if i & 1 else 0] += K[2, s]
                            if j > 0:
                                co[0 if j & 1 else 3] += K[3, s]
                            if j < ny - 1:
                                co[4 if j & 1 else 0] += K[4, s]
                            if nz > 1:
                                if k > 0:
                                    co[0 if k & 1 else 5] += K[5, s]
                                if k < nz - 1:
                                    co[6 if k & 1 else 0] += K[6, s]
                Kc[:, idx(cd, ic, jc, kc)] = co
    return Kc, cd


def hierarchy(A, dims):
    """Levels until at most COARSE cells (the allocation rule of the hierarchy)."""
    levels = [(A, dims)]
    while A.shape[0] > COARSE and any(d > 1 for d in dims):
        Kc, dims = galerkin_rule(stencil(A, dims), dims)
        A = dense(Kc, dims)
        levels.append((A, dims))
    return levels


def colours(dims):
    nx, ny, nz = dims
    col = np.zeros(nx * ny * nz, dtype=int)
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                col[idx(dims, i, j, k)] = (i + j + k) & 1
    return col == 0, col == 1


def coarsest_solve(A, b):
    n = A.shape[0]
    s = np.trace(A) / (n * n)
    return np.linalg.solve(A + s * np.ones((n, n)), b)


def vcycle(levels, b, l=0):
    A, dims = levels[l]
    if l == len(levels) - 1:
        return coarsest_solve(A, b)
    red, black = colours(dims)
    d = np.diag(A)
    x = np.zeros_like(b)
    x[red] = b[red] / d[red]                                          # red, from zero
    x[black] = (b[black] - A[np.ix_(black, red)] @ x[red]) / d[black]
    r = b - A @ x
    r[black] = 0.0                                                    # zero after the black sweep
    x += prolongation(dims) @ vcycle(levels, prolongation(dims).T @ r, l + 1)
    x[black] = (b[black] - A[np.ix_(black, red)] @ x[red]) / d[black]
    x[red] = (b[red] - A[np.ix_(red, black)] @ x[black]) / d[red]
    return x


def pcg(A, b, apply_m, tol=1e-8, maxit=2000):
    x = np.zeros_like(b)
    r = b.copy()
    z = apply_m(r)
    p = z.copy()
    rz = r @ z
    bb = b @ b
    for it in range(1, maxit + 1):
        q = A @ p
        alpha = rz / (p @ q)
        x += alpha * p
        r -= alpha * q
        if r @ r <= tol * tol * bb:
            return x, it
        z = apply_m(r)
        rz, rz_old = r @ z, rz
        p = z + (rz / rz_old) * p
    return x, maxit


GRIDS = [(5, 4, 3), (8, 8, 4), (7, 1, 1), (2, 2, 2), (9, 6, 1), (4, 4, 4), (3, 5, 7)]


@pytest.mark.parametrize("dims", GRIDS)
def test_galerkin_rule_is_ptap(dims):
    rng = np.random.default_rng(sum(dims))
    A = pressure_matrix(dims, rng)
    Kc, cd = galerkin_rule(stencil(A, dims), dims)
    P = prolongation(dims)
    want = P.T @ A @ P
    got = dense(Kc, cd)
    scale = np.abs(A).max()
    assert np.allclose(got, want, rtol=1e-13, atol=1e-12 * scale)
    assert np.allclose(got.sum(axis=1), 0.0, atol=1e-11 * scale)   # null space kept
    assert np.allclose(got, got.T)


def test_hierarchy_sizes_follow_the_allocation_rule():
    A = pressure_matrix((16, 12, 8), np.random.default_rng(3))
    dims = [lv[1] for lv in hierarchy(A, (16, 12, 8))]
    assert dims == [(16, 12, 8), (8, 6, 4), (4, 3, 2)]   # 24 cells <= COARSE
    assert np.prod(dims[-1]) <= COARSE


@pytest.mark.parametrize("dims", [(4, 4, 4), (5, 3, 2), (8, 8, 1)])
def test_coarsest_solve_gives_the_zero_sum_solution(dims):
    rng = np.random.default_rng(7)
    A = pressure_matrix(dims, rng)
    b = rng.standard_normal(A.shape[0])
    b -= b.mean()
    x = coarsest_solve(A, b)
    assert np.allclose(A @ x, b, atol=1e-10 * np.abs(b).max())
    assert abs(x.sum()) <= 1e-10 * np.abs(x).max()


@pytest.mark.parametrize("dims", [(8, 8, 8), (9, 7, 5), (16, 16, 2)])
def test_vcycle_is_symmetric_positive_definite(dims):
    rng = np.random.default_rng(11)
    A = pressure_matrix(dims, rng)
    levels = hierarchy(A, dims)
    assert len(levels) >= 2
    n = A.shape[0]
    M = np.column_stack([vcycle(levels, e) for e in np.eye(n)])
    assert np.linalg.norm(M - M.T) <= 1e-10 * np.linalg.norm(M)
    assert np.linalg.eigvalsh(0.5 * (M + M.T)).min() > 0.0


def test_cg_with_the_vcycle_beats_the_diagonal():
    dims = (16, 16, 8)
    rng = np.random.default_rng(5)
    A = pressure_matrix(dims, rng)
    levels = hierarchy(A, dims)
    b = rng.standard_normal(A.shape[0])
    b -= b.mean()   # consistent right-hand side (the engine projects it)
    x_mg, it_mg = pcg(A, b, lambda r: vcycle(levels, r))
    x_d, it_d = pcg(A, b, lambda r: r / np.diag(A))
    for x in (x_mg, x_d):
        assert np.linalg.norm(A @ x - b) <= 1e-7 * np.linalg.norm(b)
    assert it_mg * 3 <= it_d, (it_mg, it_d)
