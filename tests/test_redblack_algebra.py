"""CPU checks of the red-black Krylov storage and the reduced (Schur complement) system of
the fused stencil engine, with the index maps and formulas its kernels use
(equelle_amd/csrc/heat_fused.hip: cell_of, offdiag, k_precond_setup, k_load_rhs,
k_schur_red/k_schur_black, k_accum, k_spmv_rb, k_ilu_black/k_ilu_red), on random
diagonally dominant 5- and 7-point matrices over odd, even, 2D, 3D and degenerate grids:

* the ranks c >> 1 of each colour (i + j + k) % 2 are dense;
* the black ILU(0) pivot equals the diagonal of the Schur complement S;
* solving S x_b = g_b and recovering x_r gives A x = b;
* the full-system SpMV in red-black storage equals A x;
* the two-pass ILU(0) apply z satisfies M z = r for M = (P + L) P^-1 (P + U).
"""
import numpy as np
import pytest

DIRS = [(-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)]


class RB:
    def __init__(self, nx, ny, nz, seed):
        rng = np.random.default_rng(seed)
        self.nx, self.ny, self.nz = nx, ny, nz
        n = self.n = nx * ny * nz
        self.h = (n + 1) // 2
        self.ijk = [(c % nx, (c // nx) % ny, c // (nx * ny)) for c in range(n)]
        self.col = np.array([(i + j + k) & 1 for i, j, k in self.ijk])
        self.nq = 7 if nz > 1 else 5
        co = np.zeros((self.nq, n))
        A = np.zeros((n, n))
        for c, (i, j, k) in enumerate(self.ijk):
            s = 0.0
            for q, (di, dj, dk) in enumerate(DIRS[:self.nq - 1]):
                ii, jj, kk = i + di, j + dj, k + dk
                if 0 <= ii < nx and 0 <= jj < ny and 0 <= kk < nz:
                    v = -rng.uniform(0.1, 1.0)
                    co[q + 1, c] = v
                    A[c, ii + nx * (jj + ny * kk)] = v
                    s += abs(v)
            co[0, c] = s + rng.uniform(0.1, 1.0)
            A[c, c] = co[0, c]
        self.A = A
        self.K = np.zeros((self.nq, 2 * self.h))
        for c in range(n):
            self.K[:, self.col[c] * self.h + (c >> 1)] = co[:, c]
        self.rng = rng

    def cell_of(self, p, cl):
        c = 2 * p
        if c >= self.n:
            return -1
        if self.col[c] != cl:
            c += 1
            if c >= self.n:
                return -1
        return c

    def nbrs(self, c):
        """(coefficient index, neighbour cell) in W, E, S, N, B, T order."""
        i, j, k = self.ijk[c]
        nx, ny, nz = self.nx, self.ny, self.nz
        out = []
        if i > 0: out.append((1, c - 1))
        if i < nx - 1: out.append((2, c + 1))
        if j > 0: out.append((3, c - nx))
        if j < ny - 1: out.append((4, c + nx))
        if nz > 1:
            if k > 0: out.append((5, c - nx * ny))
            if k < nz - 1: out.append((6, c + nx * ny))
        return out

    def offdiag(self, s, c, x):
        return sum(self.K[q, s] * x[m >> 1] for q, m in self.nbrs(c))

    def dinv(self):
        h, K = self.h, self.K
        dv = np.zeros(2 * h)
        opposite = {1: 2, 2: 1, 3: 4, 4: 3, 5: 6, 6: 5}
        for p in range(h):
            c = self.cell_of(p, 0)
            if c >= 0:
                dv[p] = 1.0 / K[0, p]
            c = self.cell_of(p, 1)
            if c < 0:
                continue
            s = h + p
            piv = K[0, s]
            for q, m in self.nbrs(c):
                piv -= K[q, s] * K[opposite[q], m >> 1] / K[0, m >> 1]
            dv[s] = 1.0 / piv
        return dv


GRIDS = [(4, 3, 1), (5, 3, 1), (3, 5, 1), (5, 5, 1), (6, 4, 1), (7, 1, 1), (1, 7, 1), (3, 4, 3), (4, 3, 3),
         (3, 3, 3), (5, 4, 2), (2, 3, 5), (4, 4, 4), (1, 1, 1), (2, 1, 1)]


@pytest.mark.parametrize("dims", GRIDS)
def test_reduced_system_and_storage(dims):
    g = RB(*dims, seed=sum(dims))
    n, h = g.n, g.h
    for q in (0, 1):   # dense ranks per colour
        r = sorted(c >> 1 for c in range(n) if g.col[c] == q)
        assert r == list(range(len(r)))
    dv = g.dinv()
    b = g.rng.normal(size=n)
    # reduced right-hand side g_b = b_b - A_br C_r^-1 b_r (k_load_rhs)
    gb = np.zeros(h)
    for p in range(h):
        c = g.cell_of(p, 1)
        if c >= 0:
            gb[p] = b[c] - sum(g.K[q, h + p] * (b[m] * dv[m >> 1]) for q, m in g.nbrs(c))

    def S(v):   # k_schur_red then k_schur_black
        y = np.zeros(h)
        for p in range(h):
            c = g.cell_of(p, 0)
            y[p] = 0.0 if c < 0 else g.offdiag(p, c, v) * dv[p]
        w = np.zeros(h)
        for p in range(h):
            c = g.cell_of(p, 1)
            w[p] = 0.0 if c < 0 else g.K[0, h + p] * v[p] - g.offdiag(h + p, c, y)
        return w

    Sm = np.column_stack([S(e) for e in np.eye(h)])
    nb = int(sum(g.col))
    for p in range(nb):   # black ILU(0) pivot = diag(S)
        assert Sm[p, p] * dv[h + p] == pytest.approx(1.0, rel=1e-12)
    xb = np.zeros(h)
    if nb:
        xb[:nb] = np.linalg.solve(Sm[:nb, :nb], gb[:nb])
    x = np.zeros(n)
    for p in range(h):   # k_accum: recover x_r
        cr = g.cell_of(p, 0)
        if cr >= 0:
            x[cr] = (b[cr] - g.offdiag(p, cr, xb)) * dv[p]
        cb = g.cell_of(p, 1)
        if cb >= 0:
            x[cb] = xb[p]
    assert np.abs(g.A @ x - b).max() < 1e-10 * max(1.0, np.abs(b).max())


@pytest.mark.parametrize("dims", GRIDS)
def test_full_system_spmv_and_ilu(dims):
    g = RB(*dims, seed=7 + sum(dims))
    n, h = g.n, g.h
    dv = g.dinv()

    def to_rb(v):
        out = np.zeros(2 * h)
        for c in range(n):
            out[g.col[c] * h + (c >> 1)] = v[c]
        return out

    xn = g.rng.normal(size=n)
    xf = to_rb(xn)
    yn = g.A @ xn
    for p in range(h):   # k_spmv_rb
        for cl in (0, 1):
            c = g.cell_of(p, cl)
            if c >= 0:
                s = cl * h + p
                y = g.K[0, s] * xf[s] + g.offdiag(s, c, xf[(1 - cl) * h:])
                assert y == pytest.approx(yn[c], rel=1e-12, abs=1e-12)
    r = to_rb(g.rng.normal(size=n))
    z = np.zeros(2 * h)
    for p in range(h):   # k_ilu_black
        c = g.cell_of(p, 1)
        if c >= 0:
            s = h + p
            z[s] = (r[s] - sum(g.K[q, s] * (r[m >> 1] * dv[m >> 1]) for q, m in g.nbrs(c))) * dv[s]
    for p in range(h):   # k_ilu_red
        c = g.cell_of(p, 0)
        if c >= 0:
            z[p] = (r[p] - g.offdiag(p, c, z[h:])) * dv[p]
    # M z = (P + L) P^-1 (P + U) z in red-black storage must give back r
    Pd = np.zeros(2 * h)
    Pd[:h] = [g.K[0, p] if g.cell_of(p, 0) >= 0 else 1.0 for p in range(h)]
    Pd[h:] = [1.0 / dv[h + p] if g.cell_of(p, 1) >= 0 else 1.0 for p in range(h)]
    t = Pd * z   # (P + U) z: U = red rows' black neighbours
    for p in range(h):
        c = g.cell_of(p, 0)
        if c >= 0:
            t[p] += g.offdiag(p, c, z[h:])
    t = t / Pd
    out = Pd * t   # (P + L) t: L = black rows' red neighbours
    for p in range(h):
        c = g.cell_of(p, 1)
        if c >= 0:
            out[h + p] += g.offdiag(h + p, c, t[:h])
    for p in range(h):
        for cl in (0, 1):
            if g.cell_of(p, cl) >= 0:
                s = cl * h + p
                assert out[s] == pytest.approx(r[s], rel=1e-10, abs=1e-12)
