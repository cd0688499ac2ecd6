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


@pytest.mark.parametrize("src", ["nonlinear_heat_cuda.cpp", "twophase_corey_cuda.cpp"])
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
