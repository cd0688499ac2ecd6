"""Drop-in boundary: code in the exact shape the Equelle CUDA printer emits
(compiler/PrintCUDABackendASTVisitor.cpp:18-77) compiles unchanged against
include/EquelleRuntimeCUDA.hpp and links to libeqhip.so; on the GPU it runs and
matches the oracle."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from equelle_amd.capi import dptr, iptr

GEN = os.path.join(ROOT, "tests", "generated")
EXE = os.path.join(GEN, "bin", "nonlinear_heat_cuda")
EXE_TP = os.path.join(GEN, "bin", "twophase_corey_cuda")
EXE_FI = os.path.join(GEN, "bin", "twophase_fully_implicit_cuda")


@pytest.mark.parametrize("src", ["nonlinear_heat_cuda.cpp", "twophase_corey_cuda.cpp", "heateq_cuda.cpp",
                                 "exp_heateq_cuda.cpp", "twophase_fully_implicit_cuda.cpp",
                                 "residual_forms_cuda.cpp"])
def test_generated_program_compiles(tmp_path, src):
    exe = tmp_path / "prog"
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"),
                           os.path.join(GEN, src),
                           "-L" + os.path.join(ROOT, "equelle_amd", "lib"), "-leqhip", "-o", str(exe)])
    assert exe.exists()


def _write_inputs(d, nx, ny, u0, dirf, dirv, dts):
    (d / "u_initial.txt").write_text("\n".join(f"{x:.17g}" for x in u0))
    (d / "dir_boundary.txt").write_text(" ".join(str(int(f)) for f in dirf))
    (d / "dir_val.txt").write_text(" ".join(f"{x:.17g}" for x in dirv))
    (d / "timesteps.txt").write_text(" ".join(f"{x:.17g}" for x in dts))
    (d / "run.param").write_text("\n".join([
        "grid_dim=2", f"nx={nx}", f"ny={ny}", "dx=1", "dy=1", "k=0.1", "cv=1",
        "u_initial_from_file=true", "u_initial_filename=u_initial.txt",
        "dir_boundary_filename=dir_boundary.txt",
        "dir_val_from_file=true", "dir_val_filename=dir_val.txt",
        "timesteps_filename=timesteps.txt", "output_to_file=true",
        "abs_res_tol=0", "rel_res_tol=1e-11", "max_iter=25", "solver_tol=1e-12"]) + "\n")


@pytest.mark.gpu
def test_generated_program_runs_and_matches_oracle(tmp_path, oracle_api):
    assert os.path.exists(EXE), "run __graft_entry__.build() first"
    nx, ny = 48, 40
    rng = np.random.default_rng(9)
    u0 = rng.uniform(0, 2, nx * ny)
    dirf = np.array([j * (nx + 1) for j in range(ny)] + [nx + j * (nx + 1) for j in range(ny)], dtype=np.int32)
    dirf.sort()
    dirv = np.where(dirf % (nx + 1) == 0, 0.0, 1.0)
    dts = np.array([0.5, 1.0, 2.0])
    _write_inputs(tmp_path, nx, ny, u0, dirf, dirv, dts)
    subprocess.run([EXE, "run.param"], cwd=tmp_path, check=True, timeout=120, stdout=subprocess.DEVNULL)
    u_last = np.loadtxt(tmp_path / "u-00002.output")
    params = (f"grid_dim=2 nx={nx} ny={ny} k=0.1 cv=1 k_model=quadratic abs_res_tol=0 rel_res_tol=1e-11 "
              "max_iter=25").encode()
    uo = np.zeros(nx * ny)
    oracle_api.heat_run(params, dptr(u0), nx * ny, iptr(dirf), dptr(dirv), len(dirf), dptr(dts), 3, dptr(uo),
                        None, None)
    assert np.allclose(u_last, uo, rtol=1e-9, atol=1e-10)


@pytest.mark.gpu
def test_generated_twophase_runs_and_matches_oracle(tmp_path, oracle_api):
    """The two-phase program (3D, file inputs, InputDomainSubsetOf, Centroid/Normal/Dot,
    two NewtonSolves per step, Output) against the oracle's run of the same program."""
    from twophase_cases import case, oracle_run, twophase_params
    assert os.path.exists(EXE_TP), "run __graft_entry__.build() first"
    nx, ny, nz = 10, 8, 5
    perm, poro, sw0, sc, sv = case(nx, ny, nz)
    dts = np.array([1.0, 1.0, 2.0])
    d = tmp_path
    for name, arr in (("perm", perm), ("poro", poro), ("sw_initial", sw0), ("source_values", sv), ("timesteps", dts)):
        (d / f"{name}.txt").write_text("\n".join(f"{x:.17g}" for x in arr))
    (d / "source_cells.txt").write_text(" ".join(str(int(c)) for c in sc))
    (d / "run.param").write_text("\n".join([
        "grid_dim=3", f"nx={nx}", f"ny={ny}", f"nz={nz}", "dx=1", "dy=1", "dz=1",
        "perm_from_file=true", "perm_filename=perm.txt", "poro_from_file=true", "poro_filename=poro.txt",
        "sw_initial_from_file=true", "sw_initial_filename=sw_initial.txt",
        "source_cells_filename=source_cells.txt",
        "source_values_from_file=true", "source_values_filename=source_values.txt",
        "timesteps_filename=timesteps.txt", "output_to_file=true", "solver=GMRes", "preconditioner=diagonal",
        "abs_res_tol=0", "rel_res_tol=1e-12", "max_iter=40", "solver_tol=1e-13"]) + "\n")
    subprocess.run([EXE_TP, "run.param"], cwd=tmp_path, check=True, timeout=120, stdout=subprocess.DEVNULL)
    sw_last = np.loadtxt(tmp_path / "saturation-00002.output")
    p_last = np.loadtxt(tmp_path / "pressure-00002.output")
    params = twophase_params(nx, ny, nz, rel_res_tol=1e-13, max_iter=40).encode()
    po, so, _ = oracle_run(oracle_api, params, perm, poro, sw0, sc, sv, dts)
    assert np.allclose(sw_last, so, rtol=1e-9, atol=1e-10)
    assert np.allclose(p_last - p_last[-1], po - po[-1], rtol=1e-9, atol=1e-9 * np.abs(po).max())


@pytest.mark.gpu
def test_generated_fully_implicit_runs_and_matches_oracle(tmp_path, oracle_api):
    """The fully implicit two-phase program (twophase_fully_implicit.equelle: one
    NewtonSolveSystem on (p, sw) per step, makeArray of two lambdas) compiled against the
    facade and run, against the oracle's run of the same program."""
    from twophase_cases import case, fi_params, oracle_fi_run
    assert os.path.exists(EXE_FI), "run __graft_entry__.build() first"
    nx, ny, nz = 7, 6, 4
    perm, poro, sw0, sc, sv = case(nx, ny, nz)
    dts = np.array([0.05, 0.05, 0.1])   # steps plain Newton converges on from the guess sw = 0.5
    d = tmp_path
    for name, arr in (("perm", perm), ("poro", poro), ("sw_initial", sw0), ("source_values", sv), ("timesteps", dts)):
        (d / f"{name}.txt").write_text("\n".join(f"{x:.17g}" for x in arr))
    (d / "source_cells.txt").write_text(" ".join(str(int(c)) for c in sc))
    (d / "run.param").write_text("\n".join([
        "grid_dim=3", f"nx={nx}", f"ny={ny}", f"nz={nz}", "dx=1", "dy=1", "dz=1", "watervisc=0.001",
        "perm_from_file=true", "perm_filename=perm.txt", "poro_from_file=true", "poro_filename=poro.txt",
        "sw_initial_from_file=true", "sw_initial_filename=sw_initial.txt",
        "source_cells_filename=source_cells.txt",
        "source_values_from_file=true", "source_values_filename=source_values.txt",
        "timesteps_filename=timesteps.txt", "output_to_file=true", "solver=GMRes", "preconditioner=diagonal",
        "abs_res_tol=0", "rel_res_tol=1e-12", "max_iter=40", "solver_tol=1e-7"]) + "\n")
    subprocess.run([EXE_FI, "run.param"], cwd=tmp_path, check=True, timeout=120, stdout=subprocess.DEVNULL)
    sw_last = np.loadtxt(tmp_path / "saturation-00003.output")   # 00000 is the initial state
    p_last = np.loadtxt(tmp_path / "pressure-00003.output")
    params = fi_params(nx, ny, nz, rel_res_tol=1e-13, max_iter=40).encode()
    po, so, _ = oracle_fi_run(oracle_api, params, perm, poro, sw0, sc, sv, dts)
    assert np.allclose(sw_last, so, rtol=1e-9, atol=1e-10)
    assert np.allclose(p_last - p_last[-1], po - po[-1], rtol=1e-9, atol=1e-9 * np.abs(po).max())


EXE_HEATEQ = os.path.join(GEN, "bin", "heateq_cuda")
EXE_EXP = os.path.join(GEN, "bin", "exp_heateq_cuda")
WEBDEMO = os.path.join(ROOT, "tests", "golden", "webdemo_2dheateq")


def _webdemo_params():
    """The parameter file the web demo's runner writes for webdemo/srv/examples/2dheateq
    (config.json: 2D, 100 x 100 cells of 10 x 10, abs_res_tol 1e-4; one `<tag>_from_file` /
    `<tag>_filename=../<tag>` pair per input file; executableRun.js:26-61), in its order."""
    lines = ["grid_dim=2", "ny=100", "dy=10", "nx=100", "dx=10", "abs_res_tol=0.0001"]
    for tag in ("timesteps", "dir_boundary", "dir_val", "u_initial"):
        lines += [f"{tag}_from_file=true", f"{tag}_filename=../{tag}"]
    return "\n".join(lines + ["output_to_file=true", "verbose=2"]) + "\n"


@pytest.mark.gpu
def test_generated_heateq_runs_the_webdemo_inputs(tmp_path, oracle_api):
    """examples/dsl/heateq.equelle in printer form on the reference's own web-demo inputs,
    unchanged (the four files of webdemo/srv/examples/2dheateq under their input tags),
    with the parameter file the demo writes.  Every u-NNNNN.output (the initial field and
    ten steps of dt = 1e8 s; granite defaults k = 2.85, cv = 2.17e6) against the oracle's
    run of the same program solved to 1e-12.  Bar: the program stops at |F| <= 1e-4 after
    a BiCGStab solve to 1e-8 relative, one Newton iteration per step (the problem is
    linear); with ||J^-1|| <= 1 the fields may differ by 1e-8 ||F0|| per step, so 1e-6 of
    the largest temperature."""
    import shutil
    assert os.path.exists(EXE_HEATEQ), "run __graft_entry__.build() first"
    files = {"timesteps": "timesteps.mat", "dir_boundary": "dir_boundary", "dir_val": "dir_val",
             "u_initial": "u_initial"}
    for tag, name in files.items():
        shutil.copyfile(os.path.join(WEBDEMO, name), tmp_path / tag)
    (tmp_path / "params.param").write_text(_webdemo_params())
    run = tmp_path / "outputs"
    run.mkdir()
    out = subprocess.run([EXE_HEATEQ, "../params.param"], cwd=run, check=True, timeout=300, capture_output=True,
                         text=True).stdout
    assert out.count("Newton solver converged") == 10, out[-2000:]
    u0 = np.loadtxt(os.path.join(WEBDEMO, "u_initial"))
    dirf = np.loadtxt(os.path.join(WEBDEMO, "dir_boundary"), dtype=np.int32)
    dirv = np.loadtxt(os.path.join(WEBDEMO, "dir_val"))
    dts = np.loadtxt(os.path.join(WEBDEMO, "timesteps.mat"))
    assert np.array_equal(np.loadtxt(run / "u-00000.output"), u0)
    params = ("grid_dim=2 nx=100 ny=100 dx=10 dy=10 k=2.85 cv=2170000 k_model=const abs_res_tol=0 "
              "rel_res_tol=1e-12 max_iter=10").encode()
    uo = u0.copy()
    for s, dt in enumerate(dts):
        nxt = np.zeros_like(uo)
        oracle_api.heat_run(params, dptr(uo), uo.size, iptr(dirf), dptr(dirv), len(dirf), dptr(np.array([dt])), 1,
                            dptr(nxt), None, None)
        uo = nxt
        got = np.loadtxt(run / f"u-{s + 1:05d}.output")
        assert np.abs(got - uo).max() <= 1e-6 * np.abs(uo).max(), (s, np.abs(got - uo).max())
    maxima = [float(line.split("=")[1]) for line in out.splitlines() if line.startswith("maximum of u = ")]
    assert len(maxima) == 11 and maxima[-1] == pytest.approx(uo.max(), rel=1e-5)   # printed to 6 digits


def _simulator_inputs(d, nx, ny, nz, timesteps, dt):
    """The input files of examples/simulators/heateq/createDirichlet.m for an nx x ny x nz
    grid: timesteps.mat (timesteps x dt), dir_bnd.mat (both x-faces of every row of
    cells, ascending), dir_val.mat (per (y, z): a value falling linearly from 1000 to 0 over
    y at x = 0, its mirror in y at x = max, listed y-major as the script writes them)."""
    (d / "timesteps.mat").write_text(" ".join(f"{dt:.7e}" for _ in range(timesteps)) + "\n")
    idx = np.zeros(2 * ny * nz, dtype=np.int64)
    for i in range(1, ny * nz):
        idx[2 * i - 1] = i * (nx + 1) - 1
        idx[2 * i] = i * (nx + 1)
    idx[-1] = ny * nz * (nx + 1) - 1
    (d / "dir_bnd.mat").write_text("".join(f"{v}\n" for v in idx))
    val_y = np.linspace(1000.0, 0.0, ny)
    vals = np.zeros(2 * ny * nz)
    for y in range(ny):
        for z in range(nz):
            vals[2 * (y * nz + z)] = val_y[y]
            vals[2 * (y * nz + z) + 1] = val_y[ny - 1 - y]
    (d / "dir_val.mat").write_text("".join(f"{v:.7e}\n" for v in vals))
    return idx.astype(np.int32), np.array([float(f"{v:.7e}") for v in vals]), np.full(timesteps, float(f"{dt:.7e}"))


@pytest.mark.gpu
def test_generated_exp_heateq_runs_the_simulator_params(tmp_path, oracle_api):
    """examples/simulators/heateq/exp_heateq.equelle (explicit heat with Dirichlet faces) in
    printer form, run with that simulator's params.param keys unchanged except the grid
    size, which performanceTest.sh also rewrites per case, and the inputs its
    createDirichlet.m writes for that size (120 steps of dt = 0.5, uniform u_initial = 500).
    Every expU output against the same program evaluated on the oracle."""
    from equelle_amd.runtime import EquelleRuntimeCUDA
    assert os.path.exists(EXE_EXP), "run __graft_entry__.build() first"
    nx, ny, nz, nsteps = 20, 24, 10, 120
    dirf, dirv, dts = _simulator_inputs(tmp_path, nx, ny, nz, nsteps, 0.5)
    (tmp_path / "params.param").write_text("\n".join([
        "grid_dim=3", f"nx={nx}", "dx=1", f"ny={ny}", "dy=1", f"nz={nz}", "dz=1", "k=0.3",
        "u_initial_from_file=false", "u_initial=500",
        "dirichlet_boundary_from_file=true", "dirichlet_boundary_filename=../dir_bnd.mat",
        "dirichlet_val_from_file=true", "dirichlet_val_filename=../dir_val.mat",
        "timesteps_from_file=true", "timesteps_filename=../timesteps.mat",
        "output_to_file=true", "solver=CG", "preconditioner=diagonal"]))
    run = tmp_path / "run"
    run.mkdir()
    subprocess.run([EXE_EXP, "../params.param"], cwd=run, check=True, timeout=300, stdout=subprocess.DEVNULL)
    er = EquelleRuntimeCUDA(f"grid_dim=3 nx={nx} ny={ny} nz={nz} dx=1 dy=1 dz=1", api=oracle_api)
    allc, allf, bf, ifs = er.allCells(), er.allFaces(), er.boundaryFaces(), er.interiorFaces()
    db = er.indices(dirf, 1)
    dval = er.collOfScalar(dirv)
    vol = er.norm(allc)
    first, second = er.firstCell(ifs), er.secondCell(ifs)
    itrans = 0.3 * (er.norm(ifs) / er.norm(er.centroid(first) - er.centroid(second)))
    emp = er.isEmpty(er.firstCell(bf))
    bf_cells = er.trinaryIf(emp, er.secondCell(bf), er.firstCell(bf))
    bf_sign = er.trinaryIf(emp, er.operatorExtend(-1.0, bf), er.operatorExtend(1.0, bf))
    btrans = 0.3 * (er.norm(bf) / er.norm(er.centroid(bf) - er.centroid(bf_cells)))
    dir_sign = er.operatorOn(bf_sign, bf, db)
    u = er.collOfScalar(np.full(nx * ny * nz, 500.0))
    for s, dt in enumerate(dts):
        udir = er.operatorOn(u, allc, er.operatorOn(bf_cells, bf, db))
        dflux = (er.operatorOn(btrans, bf, db) * dir_sign) * (udir - dval)
        fluxes = er.operatorExtend((-itrans) * er.gradient(u), ifs, allf) + \
            er.operatorExtend(er.operatorExtend(dflux, db, bf), bf, allf)
        u = u - (dt / vol) * er.divergence(fluxes)
        if s % 20 == 19 or s == 0:
            got = np.loadtxt(run / f"expU-{s:05d}.output")
            assert np.allclose(got, u.value(), rtol=1e-12, atol=1e-9), (s, np.abs(got - u.value()).max())
    assert np.allclose(np.loadtxt(run / f"expU-{nsteps:05d}.output"), u.value(), rtol=1e-12, atol=1e-9)


EXE_FORMS = os.path.join(GEN, "bin", "residual_forms_cuda")


def _forms_residual(form, u, uprev, dt, nx, ny, k=0.1):
    """The residuals of residual_forms_cuda.cpp on the unit-spaced nx x ny Cartesian grid,
    restated in numpy from the program text (face lists of the grid built here: x faces
    between cells i and i+1, y faces between rows j and j+1; the first cell is the lower
    one, so the gradient is u[second] - u[first] and the transmissibility is 1)."""
    c = np.arange(nx * ny).reshape(ny, nx)
    first = np.concatenate([c[:, :-1].ravel(), c[:-1, :].ravel()])
    second = np.concatenate([c[:, 1:].ravel(), c[1:, :].ravel()])

    def div(flux):
        return np.bincount(first, flux, nx * ny) - np.bincount(second, flux, nx * ny)

    if form == 0:
        kc = k * (1 + np.sqrt(u * u + 0.25))
    elif form == 1:
        kc = k * (1 + u * u)
    else:
        kc = k / (1 + u.max()) * (1 + u * u)
    kf = 0.5 * (kc[first] + kc[second])
    flux = -kf * (u[second] - u[first])
    sink = np.zeros_like(u)
    if form == 0:
        up = np.where(u[first] > u[second], u[first], u[second])
        flux = flux * (1 + 0.1 * np.abs(up)) / (1 + 0.05 * up * up)
        bnd = np.zeros((ny, nx), dtype=bool)
        bnd[0, :] = bnd[-1, :] = bnd[:, 0] = bnd[:, -1] = True
        sink = np.where(bnd.ravel(), 0.01 * u, 0.0)
    return (u - uprev) + dt * (div(flux) + sink)


@pytest.mark.gpu
def test_generated_residual_forms_fused_and_op_by_op(tmp_path):
    """Recorded residual programs behind the facade (residual_forms_cuda.cpp): a residual
    with sqrt, norm, division, a trinaryIf on a comparison of the unknown and On / Extend of
    the boundary cells; one with a values-only coefficient (u.value()); and one with a
    reduction of the unknown inside the functor, which stops the recording.  The fused run
    (fuse_residual=1) and the op-by-op run (fuse_residual=0, the reference's evaluation)
    agree to 1e-10, and every saved state solves its step's equation, evaluated in numpy."""
    assert os.path.exists(EXE_FORMS), "run __graft_entry__.build() first"
    nx, ny = 40, 32
    n = nx * ny
    outs, logs = {}, {}
    for fuse in (1, 0):
        d = tmp_path / f"fuse{fuse}"
        d.mkdir()
        (d / "run.param").write_text("\n".join([
            "grid_dim=2", f"nx={nx}", f"ny={ny}", "dx=1", "dy=1", "k=0.1", "output_to_file=true",
            "abs_res_tol=0", "rel_res_tol=1e-12", "max_iter=80", "solver_tol=1e-13",
            f"fuse_residual={fuse}"]) + "\n")
        r = subprocess.run([EXE_FORMS, "run.param"], cwd=d, check=True, timeout=120, capture_output=True,
                           text=True)
        logs[fuse] = {}
        for line in r.stdout.splitlines():
            if " stats " in line:
                head, rest = line.split(" stats ")
                logs[fuse][int(head.split()[1])] = {kv.split("=")[0]: int(kv.split("=")[1]) for kv in rest.split()}
        outs[fuse] = {(f, s): np.loadtxt(d / f"form{f}-{s:05d}.output") for f in range(3) for s in range(3)}
    s = logs[1]
    assert set(s) == {0, 1, 2}, s
    assert s[2]["rejected"] == 0, s
    # forms 0 and 1 run fused: recorded once, then found symbolically on the later steps
    assert s[1]["aborted"] == 0 and s[0]["evaluations"] > 0 and s[1]["evaluations"] > s[0]["evaluations"], s
    assert s[0]["symbolic"] >= 2 and s[1]["symbolic"] >= s[0]["symbolic"] + 2, s
    # form 2 reads a reduction of the unknown: its recording stops, and it runs op by op
    assert s[2]["aborted"] >= 1 and s[2]["evaluations"] == s[1]["evaluations"], s
    assert all(v["evaluations"] == 0 for v in logs[0].values()), logs[0]
    u_init = np.array([0.5 + 0.4 * np.sin(0.37 * c) + 0.3 * np.cos(0.11 * c) for c in range(n)])
    dts = [0.5, 1.0, 1.0]
    for f in range(3):
        prev = u_init
        for st in range(3):
            uf, ue = outs[1][(f, st)], outs[0][(f, st)]
            assert np.allclose(uf, ue, rtol=1e-10, atol=1e-10), (f, st, np.abs(uf - ue).max())
            res = _forms_residual(f, uf, prev, dts[st], nx, ny)
            assert np.abs(res).max() < 1e-9, (f, st, np.abs(res).max())
            prev = uf
