"""GPU parity: the HIP runtime (through its C-ABI) against the CPU oracle and the
reference's known answers.

Bars: grid topology, index sets and CSR sparsity patterns bit-exact; FP64 values
within 1e-10 relative (atol 1e-12 scaled to the data); Krylov results through the
residual they leave (||b - A x|| <= tol ||b||) and Newton solutions within the
tolerance the solve was run to.
"""
import ctypes as C

import numpy as np
import pytest

from equelle_amd.capi import dptr, iptr
from equelle_amd.runtime import EquelleRuntimeCUDA
from equelle_cases import (c0_dirichlet, csr_dense, heat_program, known_answers,
                           vals_on_grid_params, vals_on_grid_sequence)

pytestmark = pytest.mark.gpu
KA = known_answers()
RT = 1e-10


def pair(hip_api, oracle_api, params):
    return EquelleRuntimeCUDA(params, api=hip_api), EquelleRuntimeCUDA(params, api=oracle_api)


def close(a, b, rtol=RT):
    a, b = np.asarray(a), np.asarray(b)
    scale = max(1.0, float(np.abs(b).max())) if b.size else 1.0
    return a.shape == b.shape and np.allclose(a, b, rtol=rtol, atol=1e-12 * scale)


def same_jac(x, y, rtol=RT):
    dx, dy = x.derivative(), y.derivative()
    if dx is None or dy is None:
        return dx is None and dy is None
    rpx, cix, vx, rx, cx = dx
    rpy, ciy, vy, ry, cy = dy
    return (rx, cx) == (ry, cy) and np.array_equal(rpx, rpy) and np.array_equal(cix, ciy) and close(vx, vy, rtol)


# ------------------------------------------------------------------ topology / geometry
def test_grid_4x3_known_answers(hip_api):
    g = KA["grid4x3"]
    er = EquelleRuntimeCUDA(g["params"], api=hip_api)
    assert er.boundaryFaces().stdToHost().tolist() == g["boundary_faces"]
    assert er.interiorFaces().stdToHost().tolist() == g["interior_faces"]
    assert er.boundaryCells().stdToHost().tolist() == g["boundary_cells"]
    assert er.interiorCells().stdToHost().tolist() == g["interior_cells"]
    assert er.firstCell(er.allFaces()).stdToHost().tolist() == g["first_cell_all"]
    assert er.secondCell(er.allFaces()).stdToHost().tolist() == g["second_cell_all"]
    fb, sb = er.firstCell(er.boundaryFaces()), er.secondCell(er.boundaryFaces())
    assert fb.stdToHost().tolist() == g["first_cell_boundary"]
    assert sb.stdToHost().tolist() == g["second_cell_boundary"]
    assert er.trinaryIf(er.isEmpty(sb), fb, sb).stdToHost().tolist() == g["trinary_inner_bnd"]


@pytest.mark.parametrize("params", ["grid_dim=2 nx=37 ny=23 dx=0.7 dy=1.9",
                                    "grid_dim=3 nx=14 ny=22 nz=9 dx=1 dy=1 dz=1",
                                    "grid_dim=3 nx=5 ny=3 nz=4 dx=2 dy=0.5 dz=3"])
def test_grid_arrays_match_oracle(hip_api, oracle_api, params):
    h, o = pair(hip_api, oracle_api, params)
    assert (h.dimensions, h.nc, h.nf, h.nif, h.nbf) == (o.dimensions, o.nc, o.nf, o.nif, o.nbf)
    for w in range(3):
        assert np.array_equal(h.gridIntArray(w), o.gridIntArray(w))
    for w in range(5):
        assert close(h.gridDoubleArray(w), o.gridDoubleArray(w))
    for name in ("allCells", "boundaryCells", "interiorCells", "allFaces", "boundaryFaces", "interiorFaces"):
        assert np.array_equal(getattr(h, name)().stdToHost(), getattr(o, name)().stdToHost()), name
    for w in range(3):
        a, b = h._opmat(w), o._opmat(w)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
    for s in ("allFaces", "interiorFaces", "boundaryFaces"):
        assert close(h.normal(getattr(h, s)()).data, o.normal(getattr(o, s)()).data)
        assert close(h.centroid(getattr(h, s)()).data, o.centroid(getattr(o, s)()).data)


@pytest.mark.parametrize("which,key", [(0, "grad"), (1, "div"), (2, "fulldiv")])
def test_helper_ops_known_answers(hip_api, which, key):
    m = KA["matrix"][key]
    er = EquelleRuntimeCUDA(KA["matrix"]["params"], api=hip_api)
    rp, ci, v, rows, cols = er._opmat(which)
    assert (rows, cols) == (m["rows"], m["cols"])
    assert rp.tolist() == m["rp"] and ci.tolist() == m["ci"] and np.allclose(v, m["v"])


def test_vals_on_grid_known_answers(hip_api):
    G = KA["valsOnGrid"]
    er = EquelleRuntimeCUDA(vals_on_grid_params(), api=hip_api)
    out = vals_on_grid_sequence(er)
    for key, val in out.items():
        if key in G:
            assert np.allclose(val, G[key], rtol=1e-12, atol=1e-12), key
    assert out["subset_face"].tolist() == G["subset"]


# ------------------------------------------------------------------ CSR algebra
def _csr(api, op, A, B):
    info = np.zeros(3, dtype=np.int32)
    rp = np.zeros(A["rows"] + 1, dtype=np.int32)
    ci = np.zeros(64, dtype=np.int32)
    v = np.zeros(64)
    arr = lambda x, t: np.ascontiguousarray(x, dtype=t)
    a = [arr(A["rp"], np.int32), arr(A["ci"], np.int32), arr(A["v"], np.float64)]
    b = [arr(B["rp"], np.int32), arr(B["ci"], np.int32), arr(B["v"], np.float64)]
    api.csr_binop(op, A["rows"], A["cols"], iptr(a[0]), iptr(a[1]), dptr(a[2]), B["rows"], B["cols"],
                  iptr(b[0]), iptr(b[1]), dptr(b[2]), iptr(info), iptr(rp), iptr(ci), dptr(v), 64)
    return rp.tolist(), ci[:info[2]].tolist(), v[:info[2]]


def test_csr_algebra_known_answers(hip_api):
    M = KA["matrix"]
    for op, key in ((0, "D_B_plus_C"), (1, "F_B_minus_C"), (2, "E_B_times_C")):
        rp, ci, v = _csr(hip_api, op, M["B"], M["C"])
        assert rp == M[key]["rp"] and ci == M[key]["ci"], key
        assert np.allclose(v, M[key]["v"], atol=1e-14), key
    F = M["F_B_minus_C"]
    for tr, key in ((0, "F_times_x"), (1, "Ft_times_x")):
        y = np.zeros(4)
        hip_api.csr_spmv(4, 4, iptr(np.array(F["rp"], dtype=np.int32)), iptr(np.array(F["ci"], dtype=np.int32)),
                         dptr(np.array(F["v"], dtype=np.float64)), dptr(np.array(M[key]["x"], dtype=np.float64)),
                         dptr(y), tr)
        assert np.allclose(y, M[key]["y"], atol=1e-14)


# ------------------------------------------------------------------ AD engine
def seeded_variable(er, n):
    """autoDiff.cpp:250-265: a primary variable pushed through div.grad.div.grad so
    the Jacobian is a wide non-trivial stencil."""
    init = np.array([i + (i - (313 % 17)) * 0.1 for i in range(n)])
    u = er.primaryVariable(er.collOfScalar(init))
    g = er.gradient(u)
    x = er.divergence(g)
    g = er.gradient(x)
    return er.divergence(g)


def autodiff_sequence(er):
    """The op sequence checked by autoDiff.cpp:293-563, returning every result."""
    n = er.nc
    x = seeded_variable(er, n)
    c = er.collOfScalar(np.array([(i % 30) * 0.1 for i in range(n)]))
    pos = er.collOfScalar(np.array([1.0 + (i % 7) * 0.25 for i in range(n)]))
    allc, ifs, bf = er.allCells(), er.interiorFaces(), er.boundaryFaces()
    bc = er.boundaryCells()
    r = {}
    r["ad+const"] = x + c
    r["const+ad"] = c + x
    r["ad+ad"] = x + x
    r["ad-const"] = x - c
    r["const-ad"] = c - x
    r["ad-ad"] = x - x
    r["ad*const"] = x * c
    r["ad*ad"] = x * x
    r["ad/const"] = x / pos
    r["const/ad"] = pos / (x * x + pos)
    r["ad/ad"] = x / (x * x + pos)
    r["s*ad"] = 3.5 * x
    r["ad*s"] = x * -2.0
    r["ad/s"] = x / 4.0
    r["s/ad"] = 2.0 / (x * x + pos)
    r["-ad"] = -x
    r["on_bc"] = er.operatorOn(x, allc, bc)
    r["ext_bc"] = er.operatorExtend(r["on_bc"], bc, allc)
    r["grad"] = er.gradient(x)
    r["div_if"] = er.divergence(r["grad"])
    r["div_all"] = er.divergence(er.operatorExtend(r["grad"], ifs, er.allFaces()))
    r["sqrt"] = er.sqrt(x * x + pos)
    r["trinary"] = er.trinaryIf(x > c, x * x, c - x)
    r["trinary2"] = er.trinaryIf(x >= 0.0, c, x)
    r["on_first"] = er.operatorOn(x, allc, er.firstCell(ifs))
    r["neg_grad"] = er.negGradient(x * x - c)
    r["int_div"] = er.interiorDivergence(r["grad"] * r["neg_grad"])
    r["norm"] = er.norm(x - c)
    r["norm_const"] = er.norm(c - pos)
    red = {k: getattr(er, k + "Reduce")(x) for k in ("min", "max", "sum")}
    red["twoNorm"] = er.twoNorm(x)
    red["prod"] = er.prodReduce(er.operatorOn(pos, allc, bc))
    return r, red


@pytest.mark.parametrize("params", ["grid_dim=3 nx=14 ny=22 nz=9 dx=1 dy=1 dz=1",   # autoDiff.param
                                    "grid_dim=2 nx=40 ny=31 dx=0.5 dy=0.3"])
def test_autodiff_sequence_parity(hip_api, oracle_api, params):
    h, o = pair(hip_api, oracle_api, params)
    rh, redh = autodiff_sequence(h)
    ro, redo = autodiff_sequence(o)
    for k in ro:
        assert rh[k].useAutoDiff() == ro[k].useAutoDiff(), k
        assert close(rh[k].value(), ro[k].value()), k
        assert same_jac(rh[k], ro[k]), k
    # summation order differs (tree vs serial): tolerance scales with the sum of |x|
    absx = np.abs(ro["-ad"].value())
    for k in redo:
        tol = 1e-13 * absx.sum() if k == "sum" else 1e-12 * max(1.0, abs(redo[k]))
        assert redh[k] == pytest.approx(redo[k], rel=1e-12, abs=tol), k


def test_ad_finite_difference_parity(hip_api, oracle_api):
    """The expression exercising every rule, on both backends."""
    from test_oracle_golden import _ad_expr
    h, o = pair(hip_api, oracle_api, "grid_dim=2 nx=9 ny=7 dx=0.7 dy=1.3")
    rng = np.random.default_rng(11)
    u0 = rng.uniform(0.5, 1.5, h.nc)
    cv = rng.uniform(0.5, 2.0, h.nc)
    Rh = _ad_expr(h, h.primaryVariable(h.collOfScalar(u0)), h.collOfScalar(cv))
    Ro = _ad_expr(o, o.primaryVariable(o.collOfScalar(u0)), o.collOfScalar(cv))
    assert close(Rh.value(), Ro.value()) and same_jac(Rh, Ro)


# ------------------------------------------------------------------ heat program
def _residual(api, params, u, u0, dirf, dirv, dt):
    n = len(u)
    info = np.zeros(2, dtype=np.int32)
    nd = len(dirf)
    dirf = np.ascontiguousarray(dirf, dtype=np.int32) if nd else np.zeros(1, dtype=np.int32)
    dirv = np.ascontiguousarray(dirv, dtype=np.float64) if nd else np.zeros(1)
    api.heat_residual(params.encode(), dptr(u), dptr(u0), n, iptr(dirf), dptr(dirv), nd, dt, iptr(info),
                      None, None, None, None)
    F = np.zeros(info[0])
    rp = np.zeros(info[0] + 1, dtype=np.int32)
    ci = np.zeros(info[1], dtype=np.int32)
    jv = np.zeros(info[1])
    api.heat_residual(params.encode(), dptr(u), dptr(u0), n, iptr(dirf), dptr(dirv), nd, dt, iptr(info),
                      dptr(F), iptr(rp), iptr(ci), dptr(jv))
    return F, rp, ci, jv


@pytest.mark.parametrize("nx,ny,quadratic", [(100, 100, False), (100, 100, True), (257, 131, True)])
def test_heat_residual_jacobian_parity(hip_api, oracle_api, nx, ny, quadratic):
    params = f"grid_dim=2 nx={nx} ny={ny} k=0.3 cv=1 k_model={'quadratic' if quadratic else 'const'}"
    rng = np.random.default_rng(nx + ny)
    u, u0 = rng.uniform(0, 2, nx * ny), rng.uniform(0, 2, nx * ny)
    dirf, dirv = c0_dirichlet(nx, ny) if not quadratic else ([], [])
    Fh, rph, cih, jvh = _residual(hip_api, params, u, u0, dirf, dirv, 0.5)
    Fo, rpo, cio, jvo = _residual(oracle_api, params, u, u0, dirf, dirv, 0.5)
    assert np.array_equal(rph, rpo) and np.array_equal(cih, cio)
    assert rph[-1] == nx * ny + 2 * ((nx - 1) * ny + nx * (ny - 1))    # n + 2 nif (5-point)
    assert close(Fh, Fo) and close(jvh, jvo)


def test_heat_c0_linear_ten_steps(hip_api, oracle_api):
    """Linear heat, 100x100, k=0.3, dt=0.5, 10 steps (the CPU-runnable workload)."""
    nx = ny = 100
    params = (f"grid_dim=2 nx={nx} ny={ny} k=0.3 cv=1 k_model=const abs_res_tol=1e-12 solver_tol=1e-13 "
              "solver_max_iter=2000").encode()
    dirf, dirv = c0_dirichlet(nx, ny)
    u0 = np.random.default_rng(0).uniform(0, 1, nx * ny)
    dts = np.full(10, 0.5)
    out = {}
    for name, api in (("hip", hip_api), ("oracle", oracle_api)):
        u = np.zeros(nx * ny)
        it = np.zeros(10, dtype=np.int32)
        api.heat_run(params, dptr(u0), nx * ny, iptr(dirf), dptr(dirv), len(dirf), dptr(dts), 10, dptr(u),
                     iptr(it), None)
        out[name] = (u, it)
    assert np.all(out["hip"][1] == 1) and np.all(out["oracle"][1] == 1)   # linear: one Newton step
    assert close(out["hip"][0], out["oracle"][0], rtol=1e-10)


@pytest.mark.parametrize("extra", ["solver=BiCGStab preconditioner=diagonal precision=fp64",
                                   "solver=BiCGStab preconditioner=ilu0 precision=fp64",
                                   "solver=GMRes preconditioner=ilu0 precision=fp64",
                                   "solver=BiCGStab preconditioner=ilu0 precision=mixed",
                                   "solver=GMRes preconditioner=diagonal precision=mixed"])
def test_nonlinear_heat_newton_parity(hip_api, oracle_api, extra):
    nx, ny = 96, 80
    params = (f"grid_dim=2 nx={nx} ny={ny} k=0.1 cv=1 k_model=quadratic abs_res_tol=0 rel_res_tol=1e-10 "
              f"max_iter=20 solver_tol=1e-11 {extra}").encode()
    rng = np.random.default_rng(5)
    u0 = rng.uniform(0, 2, nx * ny)
    dts = np.array([1.0, 1.0])
    e, ed = np.zeros(1, dtype=np.int32), np.zeros(1)
    out = {}
    for name, api in (("hip", hip_api), ("oracle", oracle_api)):
        u = np.zeros(nx * ny)
        api.heat_run(params, dptr(u0), nx * ny, iptr(e), dptr(ed), 0, dptr(dts), 2, dptr(u), None, None)
        out[name] = u
    assert close(out["hip"], out["oracle"], rtol=1e-8)


@pytest.mark.parametrize("grid", ["grid_dim=2 nx=96 ny=80",            # four pair ranks per lane (nx % 8 == 0)
                                  "grid_dim=2 nx=77 ny=53",            # strided Krylov kernels
                                  "grid_dim=3 nx=20 ny=14 nz=9"])      # 7-point stencil
@pytest.mark.parametrize("extra", ["preconditioner=ilu0 precision=fp64", "preconditioner=ilu0 precision=mixed",
                                   "preconditioner=diagonal precision=fp64",
                                   "preconditioner=diagonal precision=mixed"])
def test_stencil_solve_matches_csr_solve_and_oracle(hip_api, oracle_api, grid, extra):
    """A Jacobian with the grid's stencil pattern is solved on the stencil kernels (the
    coefficients gathered out of the CSR values): the Newton trajectory equals the CSR
    (SELL-64) solve's and the oracle's within the tolerance the solves ran to."""
    dims = dict(kv.split("=") for kv in grid.split())
    n = int(dims["nx"]) * int(dims["ny"]) * int(dims.get("nz", 1))
    base = (f"{grid} k=0.1 cv=1 k_model=quadratic abs_res_tol=0 rel_res_tol=1e-10 max_iter=20 "
            f"solver=BiCGStab solver_tol=1e-11 solver_max_iter=3000 inner_tol=1e-6 {extra}")
    rng = np.random.default_rng(n)
    u0 = rng.uniform(0, 2, n)
    dts = np.array([1.0, 0.5])
    e, ed = np.zeros(1, dtype=np.int32), np.zeros(1)
    out = {}
    for name, api, params in (("stencil", hip_api, base), ("csr", hip_api, base + " stencil_solver=0"),
                              ("oracle", oracle_api, base)):
        u = np.zeros(n)
        it = np.zeros(len(dts), dtype=np.int32)
        api.heat_run(params.encode(), dptr(u0), n, iptr(e), dptr(ed), 0, dptr(dts), len(dts), dptr(u), iptr(it), None)
        out[name] = (u, it)
    assert close(out["stencil"][0], out["oracle"][0], rtol=1e-8)
    assert close(out["stencil"][0], out["csr"][0], rtol=1e-8)
    assert np.array_equal(out["stencil"][1], out["csr"][1])


# ------------------------------------------------------------------ Krylov
@pytest.mark.parametrize("solver", ["BiCGStab", "CG", "GMRes"])
@pytest.mark.parametrize("precond", ["diagonal", "ilu0", "none"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_linear_solve_residual(hip_api, oracle_api, solver, precond, precision):
    nx, ny = 64, 48
    quadratic = solver != "CG"     # CG needs the symmetric (linear) Jacobian
    base = f"grid_dim=2 nx={nx} ny={ny} k=0.2 cv=1 k_model={'quadratic' if quadratic else 'const'}"
    er = EquelleRuntimeCUDA(base + f" solver={solver} preconditioner={precond} precision={precision} "
                                   "solver_tol=1e-9 solver_max_iter=3000", api=hip_api)
    res = heat_program(er, k=0.2, cv=1.0, quadratic=quadratic)
    rng = np.random.default_rng(2)
    u = er.collOfScalar(rng.uniform(0, 2, er.nc))
    u0 = er.collOfScalar(rng.uniform(0, 2, er.nc))
    R = res(er.primaryVariable(u), u0, 2.0)
    du, iters = er.solveForUpdate(R)
    rp, ci, v, rows, cols = R.derivative()
    A = csr_dense(rp, ci, v, rows, cols)
    b = R.value()
    rel = np.linalg.norm(b - A @ du.value()) / np.linalg.norm(b)
    assert rel <= 1e-8, (rel, iters)
    assert iters > 0


@pytest.mark.parametrize("solver", ["BiCGStab", "GMRes"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_ilu0_on_a_diagonal_jacobian(hip_api, solver, precision):
    """A residual without couplings (u*u - c): its Jacobian diag(2u) takes one colour, so
    the ILU(0) apply is the pivot division alone; the update is (u*u - c) / (2u)."""
    er = EquelleRuntimeCUDA(f"grid_dim=2 nx=33 ny=17 solver={solver} preconditioner=ilu0 precision={precision} "
                            "solver_tol=1e-12 solver_max_iter=50", api=hip_api)
    rng = np.random.default_rng(8)
    uv, cv = rng.uniform(0.5, 2.0, er.nc), rng.uniform(0.1, 3.0, er.nc)
    u = er.primaryVariable(er.collOfScalar(uv))
    du, iters = er.solveForUpdate(u * u - er.collOfScalar(cv))
    want = (uv * uv - cv) / (2.0 * uv)
    assert np.allclose(du.value(), want, rtol=1e-12 if precision == "fp64" else 1e-6, atol=0), iters
    assert 1 <= iters <= 6, iters   # the preconditioner is exact: one iteration per refinement round


@pytest.mark.parametrize("precond", ["diagonal", "ilu0"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("restart", [4, 40])
def test_gmres_many_restart_cycles(hip_api, precond, precision, restart):
    """GMRES over several restart cycles (true residual recomputed at each restart) reaches
    the tolerance, and the same solve with kernel timing on (event records between the
    launches) gives the same update bit for bit.  Restart 40: Arnoldi steps past 32 columns
    take the unfused orthogonalisation kernels."""
    nx, ny = 64, 48
    er = EquelleRuntimeCUDA(f"grid_dim=2 nx={nx} ny={ny} k=0.2 cv=1 k_model=quadratic solver=GMRes "
                            f"preconditioner={precond} precision={precision} gmres_restart={restart} "
                            "solver_tol=1e-9 solver_max_iter=3000", api=hip_api)
    res = heat_program(er, k=0.2, cv=1.0, quadratic=True)
    rng = np.random.default_rng(6)
    u = er.collOfScalar(rng.uniform(0, 2, er.nc))
    u0 = er.collOfScalar(rng.uniform(0, 2, er.nc))
    R = res(er.primaryVariable(u), u0, 2.0)
    out = {}
    try:
        for mode in (0, 1):
            hip_api.timing(mode)
            du, iters = er.solveForUpdate(R)
            out[mode] = (du.value().copy(), iters)
    finally:
        hip_api.timing(0)
        hip_api.timing(2)
    assert out[0][1] == out[1][1] and out[0][1] > 12, (out[0][1], out[1][1])
    assert np.array_equal(out[0][0], out[1][0])
    rp, ci, v, rows, cols = R.derivative()
    b = R.value()
    rel = np.linalg.norm(b - csr_dense(rp, ci, v, rows, cols) @ out[0][0]) / np.linalg.norm(b)
    assert rel <= 1e-8, rel


@pytest.mark.parametrize("fp32", [0, 1])
def test_sell_spmv_parity(hip_api, oracle_api, fp32):
    nx, ny = 123, 77
    er = EquelleRuntimeCUDA(f"grid_dim=2 nx={nx} ny={ny} k=0.3 cv=1", api=hip_api)
    res = heat_program(er, k=0.3, cv=1.0, quadratic=True)
    rng = np.random.default_rng(4)
    R = res(er.primaryVariable(er.collOfScalar(rng.uniform(0, 2, er.nc))), er.collOfScalar(rng.uniform(0, 2, er.nc)),
            1.0)
    rp, ci, v, rows, cols = R.derivative()
    x = rng.standard_normal(er.nc)
    y = np.zeros(er.nc)
    hip_api.sell_spmv(er._rt, R.h, dptr(x), dptr(y), fp32)
    yo = np.zeros(er.nc)
    oracle_api.csr_spmv(rows, cols, iptr(rp), iptr(ci), dptr(v), dptr(x), dptr(yo), 0)
    # error bound scales with sum_j |A_ij x_j| (FP32: unit roundoff 6e-8, a few terms per row)
    mag = np.zeros(er.nc)
    oracle_api.csr_spmv(rows, cols, iptr(rp), iptr(ci), dptr(np.abs(v)), dptr(np.abs(x)), dptr(mag), 0)
    tol = (1e-6 if fp32 else 1e-14) * mag + 1e-300
    assert np.all(np.abs(y - yo) <= tol)


def test_error_paths(hip_api):
    er = EquelleRuntimeCUDA("grid_dim=2 nx=6 ny=4", api=hip_api)
    from equelle_amd.capi import EquelleError
    a = er.collOfScalar(np.ones(er.nc))
    b = er.collOfScalar(np.ones(er.nf))
    with pytest.raises(EquelleError):
        a + b
    with pytest.raises(EquelleError):
        er.gradient(b)
    with pytest.raises(EquelleError):
        er.operatorOn(a, er.allFaces(), er.interiorFaces())
    with pytest.raises(EquelleError):
        EquelleRuntimeCUDA("grid_dim=4", api=hip_api)
    with pytest.raises(EquelleError):
        EquelleRuntimeCUDA("solver=CPU", api=hip_api)
