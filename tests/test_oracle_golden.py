"""Pins the CPU oracle before it is trusted as the checker.

Each oracle function is checked against (a) the reference's own known answers
(transcribed in tests/golden/reference_known_answers.json with file:line), and
(b) for what those leave open, against mathematics chosen so a plausible mistake
(dropped term, wrong sign or index, transposed operand) fails:

  grid topology / numbering   gridTest.cpp:107-171 (4x3), webdemo dir_boundary ids,
                              3D invariants (every interior face joins cells 1, nx, nx*ny apart)
  geometry                    valsOnGrid.cpp:225-350; unit normals (closed form)
  HelperOps grad/div/fulldiv  matrix.cpp:401-440; fulldiv*1 = 0 on interior cells
  CSR algebra                 matrix.cpp:185-400 (explicit zeros kept)
  On/Extend/trinaryIf/grad/div valsOnGrid.cpp:95-285 + .mat fixtures
  AD product/quotient/sqrt    central finite differences on random inputs (brute force)
  pattern conventions         u-u keeps an explicit-zero diagonal; trinaryIf unions both branches
  newtonSolve + linear solve  linear heat: one Newton step equals an independently
                              assembled direct solve; steady state = closed-form linear profile
Parity unpinned (no reference fixture, no closed form): Krylov iterates of the
reference's CUSP/UMFPACK solvers (only converged Newton outcomes are compared).
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from equelle_amd.capi import EquelleError
from equelle_amd.runtime import EquelleRuntimeCUDA
from equelle_cases import (c0_dirichlet, csr_dense, heat_program, known_answers,
                           vals_on_grid_params, vals_on_grid_sequence)

KA = known_answers()
EPS10 = 10 * np.finfo(float).eps


def mk(api, params):
    return EquelleRuntimeCUDA(params, api=api)


# ---------------------------------------------------------------- topology
def test_grid_4x3_topology(oracle_api):
    g = KA["grid4x3"]
    er = mk(oracle_api, g["params"])
    assert er.allCells().size() == 12 and er.allFaces().size() == 31
    assert er.boundaryFaces().stdToHost().tolist() == g["boundary_faces"]
    assert er.interiorFaces().stdToHost().tolist() == g["interior_faces"]
    assert er.boundaryCells().stdToHost().tolist() == g["boundary_cells"]
    assert er.interiorCells().stdToHost().tolist() == g["interior_cells"]
    assert er.firstCell(er.allFaces()).stdToHost().tolist() == g["first_cell_all"]
    assert er.secondCell(er.allFaces()).stdToHost().tolist() == g["second_cell_all"]
    fb = er.firstCell(er.boundaryFaces())
    sb = er.secondCell(er.boundaryFaces())
    assert fb.stdToHost().tolist() == g["first_cell_boundary"]
    assert sb.stdToHost().tolist() == g["second_cell_boundary"]
    inner = er.trinaryIf(er.isEmpty(sb), fb, sb)
    assert inner.stdToHost().tolist() == g["trinary_inner_bnd"]


def test_grid_100x100_webdemo_ids(oracle_api):
    d = KA["dir_boundary_100x100"]
    er = mk(oracle_api, d["params"])
    fc = er.gridIntArray(0)
    for f, cells in d["expected_face_cells"].items():
        assert fc[2 * int(f): 2 * int(f) + 2].tolist() == cells
    # the webdemo Dirichlet set is accepted as a subset of BoundaryFaces()
    ids = np.loadtxt(__import__("os").path.join(
        __import__("equelle_cases").GOLDEN, "webdemo_2dheateq", "dir_boundary")).astype(np.int32)
    sub = er.indices(ids, 1)
    bfaces = er.boundaryFaces()
    oracle_api.set_contains(er._rt, bfaces.h, sub.h)


@pytest.mark.parametrize("nx,ny,nz", [(3, 4, 5), (14, 22, 9)])
def test_grid_3d_invariants(oracle_api, nx, ny, nz):
    er = mk(oracle_api, f"grid_dim=3 nx={nx} ny={ny} nz={nz} dx=1 dy=2 dz=3")
    nc = nx * ny * nz
    nf = (nx + 1) * ny * nz + nx * (ny + 1) * nz + nx * ny * (nz + 1)
    assert (er.nc, er.nf) == (nc, nf)
    nif = (nx - 1) * ny * nz + nx * (ny - 1) * nz + nx * ny * (nz - 1)
    assert er.nif == nif
    fc = er.gridIntArray(0).reshape(-1, 2)
    inter = fc[(fc[:, 0] >= 0) & (fc[:, 1] >= 0)]
    d = inter[:, 1] - inter[:, 0]
    assert set(np.unique(d).tolist()) <= {1, nx, nx * ny}
    # each cell appears in exactly 6 faces
    cnt = np.bincount(fc[fc >= 0].ravel(), minlength=nc)
    assert (cnt == 6).all()
    # fulldiv rows: sum = 0 on interior cells, sum |.| = 6
    rp, ci, v, rows, cols = er.getFulldivMatrix()
    F = csr_dense(rp, ci, v, rows, cols)
    interior = er.interiorCells().stdToHost()
    assert np.allclose(F[interior].sum(axis=1), 0)
    assert np.allclose(np.abs(F).sum(axis=1), 6)
    # cell_faces order x-,x+,y-,y+,z-,z+: first cell of x+ face is the cell itself
    cf = er.gridIntArray(2).reshape(nc, 6)
    c = np.arange(nc)
    assert (fc[cf[:, 1], 0] == c).all() and (fc[cf[:, 0], 1] == c).all()
    assert (fc[cf[:, 5], 0] == c).all() and (fc[cf[:, 4], 1] == c).all()
    # volumes and areas (closed form)
    assert np.allclose(er.gridDoubleArray(1), 6.0)
    areas = er.gridDoubleArray(3)
    nxf, nyf = (nx + 1) * ny * nz, nx * (ny + 1) * nz
    assert np.allclose(areas[:nxf], 6.0) and np.allclose(areas[nxf:nxf + nyf], 3.0)
    assert np.allclose(areas[nxf + nyf:], 2.0)


# ---------------------------------------------------------------- HelperOps
@pytest.mark.parametrize("which,key", [(0, "grad"), (1, "div"), (2, "fulldiv")])
def test_helper_ops_4x3(oracle_api, which, key):
    m = KA["matrix"][key]
    er = mk(oracle_api, KA["matrix"]["params"])
    rp, ci, v, rows, cols = er._opmat(which)
    assert (rows, cols) == (m["rows"], m["cols"])
    assert rp.tolist() == m["rp"] and ci.tolist() == m["ci"]
    assert np.allclose(v, m["v"], atol=EPS10, rtol=0)


# ---------------------------------------------------------------- CSR algebra
def _csr(api, op, A, B):
    info = np.zeros(3, dtype=np.int32)
    cap = 64
    rp = np.zeros(A["rows"] + 1, dtype=np.int32)
    ci = np.zeros(cap, dtype=np.int32)
    v = np.zeros(cap)
    arr = lambda x, t: np.ascontiguousarray(x, dtype=t)
    a = [arr(A["rp"], np.int32), arr(A["ci"], np.int32), arr(A["v"], np.float64)]
    b = [arr(B["rp"], np.int32), arr(B["ci"], np.int32), arr(B["v"], np.float64)]
    from equelle_amd.capi import dptr, iptr
    api.csr_binop(op, A["rows"], A["cols"], iptr(a[0]), iptr(a[1]), dptr(a[2]),
                  B["rows"], B["cols"], iptr(b[0]), iptr(b[1]), dptr(b[2]),
                  iptr(info), iptr(rp), iptr(ci), dptr(v), cap)
    nnz = info[2]
    return rp.tolist(), ci[:nnz].tolist(), v[:nnz]


def test_csr_algebra_matrix_cpp(oracle_api):
    M = KA["matrix"]
    B, Cm = M["B"], M["C"]
    rp, ci, v = _csr(oracle_api, 0, B, Cm)
    assert rp == M["D_B_plus_C"]["rp"] and ci == M["D_B_plus_C"]["ci"]
    assert np.allclose(v, M["D_B_plus_C"]["v"], atol=EPS10)   # includes an explicit 0 at (0,2)
    rp, ci, v = _csr(oracle_api, 1, B, Cm)
    assert rp == M["F_B_minus_C"]["rp"] and ci == M["F_B_minus_C"]["ci"]
    assert np.allclose(v, M["F_B_minus_C"]["v"], atol=EPS10)
    rp, ci, v = _csr(oracle_api, 2, B, Cm)
    assert rp == M["E_B_times_C"]["rp"] and ci == M["E_B_times_C"]["ci"]
    assert np.allclose(v, M["E_B_times_C"]["v"], atol=EPS10)   # structural 0 kept at (3,3)
    E = dict(M["E_B_times_C"], rows=4, cols=4)
    d = M["smallDiagMult"]["d"]
    D = {"rp": [0, 1, 2, 3, 4], "ci": [0, 1, 2, 3], "v": d, "rows": 4, "cols": 4}
    rp, ci, v = _csr(oracle_api, 2, D, E)
    assert rp == E["rp"] and ci == E["ci"]
    assert np.allclose(v, M["smallDiagMult"]["v"], atol=EPS10)
    with pytest.raises(EquelleError):
        W = {"rp": [0, 2, 5, 7], "ci": [1, 2, 2, 3, 4, 0, 3], "v": [3, 4, 1, 4.5, -1, 1, 1], "rows": 3, "cols": 5}
        _csr(oracle_api, 0, B, W)


@pytest.mark.parametrize("transpose,key", [(0, "F_times_x"), (1, "Ft_times_x")])
def test_csr_spmv_matrix_cpp(oracle_api, transpose, key):
    from equelle_amd.capi import dptr, iptr
    M = KA["matrix"]
    F = M["F_B_minus_C"]
    rp = np.array(F["rp"], dtype=np.int32)
    ci = np.array(F["ci"], dtype=np.int32)
    v = np.array(F["v"], dtype=np.float64)
    x = np.array(M[key]["x"], dtype=np.float64)
    y = np.zeros(4)
    oracle_api.csr_spmv(4, 4, iptr(rp), iptr(ci), dptr(v), dptr(x), dptr(y), transpose)
    assert np.allclose(y, M[key]["y"], atol=EPS10)


# ---------------------------------------------------------------- valsOnGrid
def test_vals_on_grid(oracle_api):
    G = KA["valsOnGrid"]
    er = mk(oracle_api, vals_on_grid_params())
    out = vals_on_grid_sequence(er)
    for key in ["a", "faces", "intCell", "ext_intCell_allCell", "intFace", "ext_intFace_allFace",
                "on_allCell_intCell", "subsetVals", "extendSubsetCell_boundaryCells",
                "extendSubsetFace_interiorFaces", "boundCellVals", "onSubsetCell", "fromScalar_interiorCells",
                "vol", "area", "area_bnd", "a_big", "a_big_grad", "div_a_big_grad",
                "centroid_cells_x", "centroid_cells_y", "centroid_bnd_x", "centroid_bnd_y",
                "centroid_intface_x", "centroid_intface_y"]:
        assert np.allclose(out[key], G[key], rtol=1e-12, atol=1e-12), key
    assert np.allclose(out["b"], G["b_uniform"])
    assert out["subset_face"].tolist() == G["subset"] and out["subset_cell"].tolist() == G["subset"]
    assert out["scal1"] == 2.7182
    # stored (area-scaled) normals pinned by valsOnGrid.cpp:365-388; normal() is unit length
    nrm = er.gridDoubleArray(4).reshape(-1, 2)
    assert np.allclose(nrm[:, 0], G["stored_normals_all_x"]) and np.allclose(nrm[:, 1], G["stored_normals_all_y"])
    un = er.normal(er.allFaces()).data
    assert np.allclose(np.linalg.norm(un, axis=1), 1.0)
    assert np.allclose(un[:15], [1, 0]) and np.allclose(un[15:], [0, 1])
    sq = er.sqrt(er.collOfScalar(G["sqrt_input"])).value()
    assert np.allclose(sq, np.sqrt(G["sqrt_input"]), rtol=1e-15)


def test_input_errors(oracle_api, tmp_path):
    p = vals_on_grid_params()
    bad = tmp_path / "unsorted.txt"
    bad.write_text("3 1 2")
    p["bad_filename"] = str(bad)
    er = mk(oracle_api, p)
    with pytest.raises(EquelleError):
        er.inputDomainSubsetOf("bad", er.allCells())
    out_of_set = tmp_path / "outside.txt"
    out_of_set.write_text("5 6")   # interior cells are not boundary cells
    p["outside_filename"] = str(out_of_set)
    er = mk(oracle_api, p)
    with pytest.raises(EquelleError):
        er.inputDomainSubsetOf("outside", er.boundaryCells())


# ---------------------------------------------------------------- AD rules
def _rand_grid_runtime(api, seed=0):
    er = mk(api, "grid_dim=2 nx=5 ny=4 dx=0.7 dy=1.3")
    rng = np.random.default_rng(seed)
    return er, rng


def _ad_expr(er, u, c):
    """Exercises +, -, *, /, scalar forms, sqrt, On, Extend, gradient, divergence, trinaryIf."""
    allc, ifs, allf = er.allCells(), er.interiorFaces(), er.allFaces()
    first, second = er.firstCell(ifs), er.secondCell(ifs)
    a = u * u + c * u - (2.5 * u) / (c + u * u)
    b = er.sqrt(c * u * u + c) - 3.0 / (u * u + c)
    f = er.operatorOn(a, allc, first) * er.operatorOn(b, allc, second) + er.gradient(b) * er.gradient(a)
    sel = er.trinaryIf(f >= 0.0, f, -0.5 * f)
    return a / b + er.divergence(er.operatorExtend(sel, ifs, allf)) * (u / 4.0)


def test_ad_rules_finite_difference(oracle_api):
    er, rng = _rand_grid_runtime(oracle_api)
    n = er.nc
    u0 = rng.uniform(0.5, 1.5, n)
    c = er.collOfScalar(rng.uniform(0.5, 2.0, n))
    U = er.primaryVariable(er.collOfScalar(u0))
    R = _ad_expr(er, U, c)
    rp, ci, v, rows, cols = R.derivative()
    J = csr_dense(rp, ci, v, rows, cols)
    h = 1e-6
    Jfd = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = h
        fp = _ad_expr(er, er.collOfScalar(u0 + e), c).value()
        fm = _ad_expr(er, er.collOfScalar(u0 - e), c).value()
        Jfd[:, j] = (fp - fm) / (2 * h)
    assert np.allclose(J, Jfd, rtol=1e-6, atol=1e-6)
    # values without AD equal values with AD
    assert np.allclose(_ad_expr(er, er.collOfScalar(u0), c).value(), R.value(), rtol=1e-14)


def test_neg_gradient_interior_divergence_and_norm(oracle_api):
    """negGradient, interiorDivergence and norm(CollOfScalar) (EquelleRuntimeCPU.cpp:241-254,
    345-365) against what the mathematics fixes: ngrad = -grad entry by entry, div * x equals
    the divergence of interior fluxes, |x| with Jacobian diag(sign x) x' (finite differences)."""
    er, rng = _rand_grid_runtime(oracle_api, 2)
    n = er.nc
    u0 = rng.uniform(0.5, 1.5, n)
    U = er.primaryVariable(er.collOfScalar(u0))
    one = er.collOfScalar(np.ones(n))
    q = U * U - one
    g, ng = er.gradient(q), er.negGradient(q)
    rp, ci, v, rows, cols = g.derivative()
    nrp, nci, nv, nrows, ncols = ng.derivative()
    assert (rows, cols) == (nrows, ncols) and np.array_equal(rp, nrp) and np.array_equal(ci, nci)
    assert np.array_equal(nv, -v) and np.array_equal(ng.value(), -g.value())
    # u[first] - u[second] on each interior face
    ifs = er.interiorFaces()
    f1 = er.operatorOn(q, er.allCells(), er.firstCell(ifs)).value()
    f2 = er.operatorOn(q, er.allCells(), er.secondCell(ifs)).value()
    assert np.allclose(ng.value(), f1 - f2, rtol=0, atol=1e-15)
    flux = g * ng
    d1, d2 = er.interiorDivergence(flux), er.divergence(flux)
    assert np.array_equal(d1.value(), d2.value())
    assert all(np.array_equal(a, b) for a, b in zip(d1.derivative()[:3], d2.derivative()[:3]))
    with pytest.raises(Exception, match="InteriorFaces"):
        er.interiorDivergence(er.operatorExtend(flux, ifs, er.allFaces()))
    with pytest.raises(Exception, match="AllCells"):
        er.negGradient(flux)
    # norm: values |x| and the derivative of |x| where x stays away from 0
    shift = np.where(np.arange(n) % 3 == 0, 1.6, 0.1)   # mixes signs of u0 - shift
    s = er.collOfScalar(shift)
    N = er.norm(U - s)
    assert np.array_equal(N.value(), np.abs(u0 - shift))
    rp, ci, v, rows, cols = N.derivative()
    assert np.array_equal(rp, np.arange(n + 1)) and np.array_equal(ci, np.arange(n))
    assert np.array_equal(v, np.sign(u0 - shift))
    h = 1e-7
    for j in rng.choice(n, 5, replace=False):
        e = np.zeros(n)
        e[j] = h
        fd = (np.abs(u0 + e - shift) - np.abs(u0 - e - shift)) / (2 * h)
        assert np.allclose(csr_dense(rp, ci, v, rows, cols)[:, j], fd, atol=1e-7)
    assert er.norm(s - one).derivative() is None     # a constant stays a constant
    assert np.array_equal(er.norm(s - one).value(), np.abs(shift - 1.0))


def test_pattern_conventions(oracle_api):
    er, rng = _rand_grid_runtime(oracle_api, 1)
    n = er.nc
    U = er.primaryVariable(er.collOfScalar(rng.uniform(1, 2, n)))
    # u - u keeps the diagonal as explicit zeros (csrgeam keeps zeros)
    rp, ci, v, _, _ = (U - U).derivative()
    assert rp.tolist() == list(range(n + 1)) and ci.tolist() == list(range(n)) and np.all(v == 0)
    # constant has no Jacobian; constant*u keeps u's pattern; 0*u keeps pattern with zeros
    assert er.collOfScalar(np.ones(n)).derivative() is None
    rp, ci, v, _, _ = (0.0 * U).derivative()
    assert ci.tolist() == list(range(n)) and np.all(v == 0)
    # trinaryIf: union of both branch patterns regardless of the predicate, unselected entries 0
    ifs = er.interiorFaces()
    g = er.gradient(U)                                       # 2 entries per row
    o = er.operatorOn(U, er.allCells(), er.firstCell(ifs))   # 1 entry per row
    allfalse = g > 1e300
    t = er.trinaryIf(allfalse, o, g)
    rp, ci, v, _, _ = t.derivative()
    assert np.all(np.diff(rp) == 2)
    rp2, ci2, v2, _, _ = g.derivative()
    assert ci.tolist() == ci2.tolist() and np.allclose(v, v2)
    t = er.trinaryIf(g < 1e300, o, g)
    rp, ci, v, _, _ = t.derivative()
    assert np.all(np.diff(rp) == 2)
    first = er.firstCell(ifs).stdToHost()
    for r in range(len(first)):
        row = dict(zip(ci[rp[r]:rp[r + 1]], v[rp[r]:rp[r + 1]]))
        assert row[first[r]] == 1.0 and sum(abs(x) for x in row.values()) == 1.0
    # Extend inserts empty rows, On gathers rows
    ext = er.operatorExtend(g, ifs, er.allFaces())
    rp, _, _, rows, _ = ext.derivative()
    assert rows == er.nf and rp[-1] == 2 * er.nif
    bnd = er.boundaryFaces().stdToHost()
    assert np.all(rp[bnd + 1] - rp[bnd] == 0)


def test_reductions(oracle_api):
    er, rng = _rand_grid_runtime(oracle_api, 2)
    x = rng.uniform(-1, 1, er.nc)
    c = er.collOfScalar(x)
    assert er.sumReduce(c) == pytest.approx(x.sum(), rel=1e-14)
    assert er.prodReduce(c) == pytest.approx(np.prod(x), rel=1e-12)
    assert er.maxReduce(c) == x.max() and er.minReduce(c) == x.min()
    assert er.twoNorm(c) == pytest.approx(np.linalg.norm(x), rel=1e-14)


# ---------------------------------------------------------------- Newton / heat
def _assemble_linear_heat(nx, ny, k, dt, u0, dirf, dirv):
    """Independent textbook assembly of (I + dt L) u = u0 + dt b on unit cells, cv = 1."""
    n = nx * ny
    A = sp.lil_matrix((n, n))
    b = np.zeros(n)
    for j in range(ny):
        for i in range(nx):
            c = i + nx * j
            A[c, c] += 1.0
            for di, dj in ((1, 0), (-1, 0), (0, 1), (0, -1)):
                ii, jj = i + di, j + dj
                if 0 <= ii < nx and 0 <= jj < ny:
                    A[c, c] += dt * k
                    A[c, ii + nx * jj] -= dt * k
    for f, val in zip(dirf, dirv):
        j, i = divmod(int(f), nx + 1)
        c = 0 + nx * j if i == 0 else nx - 1 + nx * j
        A[c, c] += dt * 2 * k          # half-cell transmissibility k*area/(dx/2)
        b[c] += dt * 2 * k * val
    return sp.csr_matrix(A), u0 + b


def test_linear_heat_one_newton_step_is_direct_solve(oracle_api):
    nx, ny, k, dt = 12, 9, 0.3, 0.5
    er = mk(oracle_api, f"grid_dim=2 nx={nx} ny={ny} dx=1 dy=1 abs_res_tol=1e-12")
    dirf, dirv = c0_dirichlet(nx, ny)
    res = heat_program(er, k=k, cv=1.0, dir_faces=dirf, dir_vals=dirv)
    u0h = np.random.default_rng(0).uniform(0, 1, er.nc)
    u0 = er.collOfScalar(u0h)
    u = er.newtonSolve(lambda u: res(u, u0, dt), u0)
    st = er.last_newton
    assert st.iterations == 1 and st.final_norm < 1e-12
    A, rhs = _assemble_linear_heat(nx, ny, k, dt, u0h, dirf, dirv)
    assert np.allclose(u.value(), spla.spsolve(A.tocsc(), rhs), rtol=1e-12, atol=1e-13)
    # the Jacobian is exactly the assembled matrix, pattern n + 2*nif
    U = er.primaryVariable(u0)
    rp, ci, v, rows, cols = res(U, u0, dt).derivative()
    assert rp[-1] == er.nc + 2 * er.nif
    assert np.allclose(csr_dense(rp, ci, v, rows, cols), A.toarray(), atol=1e-14)


def test_linear_heat_steady_state_closed_form(oracle_api):
    nx, ny = 20, 3
    er = mk(oracle_api, f"grid_dim=2 nx={nx} ny={ny} dx=1 dy=1 abs_res_tol=1e-10")
    dirf, dirv = c0_dirichlet(nx, ny)
    res = heat_program(er, k=0.3, cv=1.0, dir_faces=dirf, dir_vals=dirv)
    u0 = er.collOfScalar(np.zeros(er.nc))
    u = er.newtonSolve(lambda u: res(u, u0, 1e12), u0)
    x = (np.arange(nx) + 0.5) / nx
    assert np.allclose(u.value().reshape(ny, nx), np.tile(x, (ny, 1)), atol=1e-9)


def test_nonlinear_heat_newton_converges_quadratically(oracle_api):
    nx = ny = 16
    er = mk(oracle_api, f"grid_dim=2 nx={nx} ny={ny} abs_res_tol=0 rel_res_tol=1e-12 max_iter=20")
    res = heat_program(er, k=0.1, cv=1.0, quadratic=True)
    rng = np.random.default_rng(3)
    u0h = rng.uniform(0, 2, er.nc)
    u0 = er.collOfScalar(u0h)
    norms = []
    U = er.primaryVariable(u0)
    for _ in range(4):
        r = res(U, u0, 1.0)
        norms.append(er.twoNorm(r))
        du, _ = er.solveForUpdate(r)
        U = U - du
    # quadratic convergence: r_{k+1} <= C r_k^2 with a modest constant
    assert norms[2] <= 0.1 * norms[1] ** 2 and (norms[3] <= 0.1 * norms[2] ** 2 or norms[3] < 1e-12)
    u = er.newtonSolve(lambda u: res(u, u0, 1.0), u0)
    assert er.last_newton.stop_reason == 2
    # conservation: zero-flux boundary => total heat is preserved
    assert np.sum(u.value()) == pytest.approx(np.sum(u0h), rel=1e-10)


def test_heat_program_c_entry_matches_python_sequence(oracle_api):
    """The oracle's C++ heat program (used for large sizes and the CPU baseline)
    equals the same program driven op by op through the C-ABI."""
    from equelle_amd.capi import dptr, iptr
    nx, ny = 10, 7
    params = f"grid_dim=2 nx={nx} ny={ny} k=0.1 cv=1 k_model=quadratic"
    er = mk(oracle_api, params)
    res = heat_program(er, k=0.1, cv=1.0, quadratic=True)
    rng = np.random.default_rng(5)
    uh, u0h = rng.uniform(0, 2, er.nc), rng.uniform(0, 2, er.nc)
    R = res(er.primaryVariable(er.collOfScalar(uh)), er.collOfScalar(u0h), 0.7)
    rp, ci, v, rows, cols = R.derivative()
    info = np.zeros(2, dtype=np.int32)
    e = np.zeros(0, dtype=np.int32)
    ed = np.zeros(0)
    oracle_api.heat_residual(params.encode(), dptr(uh), dptr(u0h), er.nc, iptr(e), dptr(ed), 0, 0.7, iptr(info),
                             None, None, None, None)
    F = np.zeros(info[0])
    rp2 = np.zeros(info[0] + 1, dtype=np.int32)
    ci2 = np.zeros(info[1], dtype=np.int32)
    v2 = np.zeros(info[1])
    oracle_api.heat_residual(params.encode(), dptr(uh), dptr(u0h), er.nc, iptr(e), dptr(ed), 0, 0.7, iptr(info),
                             dptr(F), iptr(rp2), iptr(ci2), dptr(v2))
    assert np.array_equal(rp, rp2) and np.array_equal(ci, ci2)
    assert np.allclose(v, v2, rtol=1e-14, atol=1e-15) and np.allclose(R.value(), F, rtol=1e-14)


def test_oracle_bicgstab_ilu0_matches_direct_solve(oracle_api):
    """The oracle's iterative solver (BiCGStab + ILU(0), taken above 40000 unknowns) is the
    checker of every large parity test: pin it against the oracle's banded direct solve
    and an independent scipy sparse LU on a 43,200-cell nonlinear-heat Jacobian (not
    symmetric: k(u) differs between the two cells of a face)."""
    nx, ny = 240, 180
    rng = np.random.default_rng(5)
    u = rng.uniform(0.0, 2.0, nx * ny)
    u0 = rng.uniform(0.0, 2.0, nx * ny)
    sol = {}
    for solver in ("direct", "bicgstab"):
        er = mk(oracle_api, f"grid_dim=2 nx={nx} ny={ny} oracle_linear_solver={solver} oracle_solver_tol=1e-13 "
                            "solver_max_iter=2000")
        res = heat_program(er, k=0.1, quadratic=True)(er.primaryVariable(er.collOfScalar(u)),
                                                       er.collOfScalar(u0), 1.0)
        du, iters = er.solveForUpdate(res)
        sol[solver] = du.value()
        rp, ci, v, rows, cols = res.derivative()
        F = res.value()
    J = sp.csr_matrix((v, ci, rp), shape=(rows, cols))
    assert abs(J - J.T).max() > 1e-3                                   # not symmetric
    ref = spla.spsolve(J.tocsc(), F)
    scale = np.linalg.norm(ref)
    assert np.linalg.norm(sol["direct"] - ref) <= 1e-12 * scale
    assert np.linalg.norm(sol["bicgstab"] - ref) <= 1e-10 * scale
    assert np.linalg.norm(J @ sol["bicgstab"] - F) <= 1e-12 * np.linalg.norm(F)
