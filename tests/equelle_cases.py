"""Equelle programs and op sequences shared by the oracle and GPU parity tests.

Each function takes an ``EquelleRuntimeCUDA`` mirror (over either backend) and
runs the same built-in calls the reference's tests or generated programs make.
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def known_answers():
    with open(os.path.join(GOLDEN, "reference_known_answers.json")) as f:
        return json.load(f)


def vals_on_grid_params():
    """valsOnGrid.param with the .mat paths pointed at tests/golden/valsOnGrid."""
    d = os.path.join(GOLDEN, "valsOnGrid")
    p = {"grid_dim": 2, "nx": 4, "dx": 0.5, "ny": 3, "dy": 0.3,
         "a_from_file": "true", "a_filename": os.path.join(d, "valsOnGrid_a.mat"),
         "b_from_file": "false", "b": 124.124,
         "faces_from_file": "true", "faces_filename": os.path.join(d, "valsOnGrid_faces.mat"),
         "scal1": 2.7182,
         "intCell_from_file": "true", "intCell_filename": os.path.join(d, "valsOnGrid_intCell.mat"),
         "intFace_from_file": "true", "intFace_filename": os.path.join(d, "valsOnGrid_intFace.mat"),
         "subset_filename": os.path.join(d, "valsOnGrid_subset.mat"),
         "subsetVals_from_file": "true", "subsetVals_filename": os.path.join(d, "valsOnGrid_subsetVals.mat"),
         "boundCellVals_from_file": "true",
         "boundCellVals_filename": os.path.join(d, "valsOnGrid_boundCellVals.mat")}
    return p


def vals_on_grid_sequence(er):
    """The op sequence of valsOnGrid.cpp:95-285; returns name -> host values."""
    out = {}
    a = er.inputCollectionOfScalar("a", er.allCells())
    out["a"] = a.value()
    out["b"] = er.inputCollectionOfScalar("b", er.allCells()).value()
    out["faces"] = er.inputCollectionOfScalar("faces", er.allFaces()).value()
    intCell = er.inputCollectionOfScalar("intCell", er.interiorCells())
    out["intCell"] = intCell.value()
    out["ext_intCell_allCell"] = er.operatorExtend(intCell, er.interiorCells(), er.allCells()).value()
    intFace = er.inputCollectionOfScalar("intFace", er.interiorFaces())
    out["intFace"] = intFace.value()
    out["ext_intFace_allFace"] = er.operatorExtend(intFace, er.interiorFaces(), er.allFaces()).value()
    out["on_allCell_intCell"] = er.operatorOn(a, er.allCells(), er.interiorCells()).value()
    subset_face = er.inputDomainSubsetOf("subset", er.interiorFaces())
    subset_cell = er.inputDomainSubsetOf("subset", er.boundaryCells())
    out["subset_face"] = subset_face.stdToHost()
    out["subset_cell"] = subset_cell.stdToHost()
    subsetVals = er.inputCollectionOfScalar("subsetVals", subset_face)
    out["subsetVals"] = subsetVals.value()
    out["extendSubsetCell_boundaryCells"] = er.operatorExtend(subsetVals, subset_cell, er.boundaryCells()).value()
    out["extendSubsetFace_interiorFaces"] = er.operatorExtend(subsetVals, subset_face, er.interiorFaces()).value()
    boundVals = er.inputCollectionOfScalar("boundCellVals", er.boundaryCells())
    out["boundCellVals"] = boundVals.value()
    out["onSubsetCell"] = er.operatorOn(boundVals, er.boundaryCells(), subset_cell).value()
    out["fromScalar_interiorCells"] = er.operatorExtend(9.124, er.interiorCells()).value()
    out["vol"] = er.norm(er.allCells()).value()
    out["area"] = er.norm(er.allFaces()).value()
    out["area_bnd"] = er.norm(er.boundaryFaces()).value()
    a_big = er.trinaryIf(a > er.operatorExtend(50, er.allCells()), a * 2, a / 2)
    out["a_big"] = a_big.value()
    g = er.gradient(a_big)
    out["a_big_grad"] = g.value()
    out["div_a_big_grad"] = er.divergence(g).value()
    cc = er.centroid(er.allCells())
    out["centroid_cells_x"], out["centroid_cells_y"] = cc[0].value(), cc[1].value()
    cb = er.centroid(er.boundaryCells())
    out["centroid_bnd_x"], out["centroid_bnd_y"] = cb[0].value(), cb[1].value()
    ci = er.centroid(er.interiorFaces())
    out["centroid_intface_x"], out["centroid_intface_y"] = ci[0].value(), ci[1].value()
    out["scal1"] = er.inputScalarWithDefault("scal1", 3.14)
    return out


def heat_program(er, k=0.3, cv=1.0, quadratic=False, dir_faces=None, dir_vals=None):
    """out_heateq.cpp:28-58 as a Python closure (plus the quadratic-k variant).

    Returns computeResidual(u, u0, dt)."""
    ifaces = er.interiorFaces()
    first = er.firstCell(ifaces)
    second = er.secondCell(ifaces)
    geo = er.norm(ifaces) / er.norm(er.centroid(first) - er.centroid(second))
    itrans = geo if quadratic else k * geo
    allc, allf = er.allCells(), er.allFaces()
    bf = er.boundaryFaces()
    have_dir = dir_faces is not None and len(dir_faces) > 0
    if have_dir:
        dir_boundary = er.indices(np.asarray(dir_faces, dtype=np.int32), 1)
        dir_val = er.collOfScalar(np.asarray(dir_vals, dtype=np.float64))
        emp = er.isEmpty(er.firstCell(bf))
        bf_cells = er.trinaryIf(emp, er.secondCell(bf), er.firstCell(bf))
        bf_sign = er.trinaryIf(emp, er.operatorExtend(-1.0, bf), er.operatorExtend(1.0, bf))
        btrans = k * (er.norm(bf) / er.norm(er.centroid(bf) - er.centroid(bf_cells)))
        dir_cells = er.operatorOn(bf_cells, er.boundaryFaces(), dir_boundary)
        dir_sign = er.operatorOn(bf_sign, er.boundaryFaces(), dir_boundary)
        dir_trans = er.operatorOn(btrans, er.boundaryFaces(), dir_boundary)
    vol = er.norm(allc)

    def interior_flux(u):
        if quadratic:
            kc = k * (er.operatorExtend(1.0, allc) + u * u)
            kf = 0.5 * (er.operatorOn(kc, allc, first) + er.operatorOn(kc, allc, second))
            return ((-itrans) * kf) * er.gradient(u)
        return (-itrans) * er.gradient(u)

    def residual(u, u0, dt):
        fluxes = er.operatorExtend(interior_flux(u), ifaces, allf)
        if have_dir:
            u_dir = er.operatorOn(u, allc, dir_cells)
            dir_fluxes = (dir_trans * dir_sign) * (u_dir - dir_val)
            bfl = er.operatorExtend(dir_fluxes, dir_boundary, er.boundaryFaces())
            fluxes = fluxes + er.operatorExtend(bfl, er.boundaryFaces(), allf)
        return (u - u0) + (dt / (cv * vol)) * er.divergence(fluxes)

    return residual


def c0_dirichlet(nx, ny):
    """Linear-heat workload boundary: u = 0 on x=0 faces, u = 1 on x=max faces."""
    faces = []
    vals = []
    for j in range(ny):
        faces.append(j * (nx + 1))
        faces.append(nx + j * (nx + 1))
        vals.extend([0.0, 1.0])
    return np.array(faces, dtype=np.int32), np.array(vals)


def csr_dense(rp, ci, v, rows, cols):
    m = np.zeros((rows, cols))
    for r in range(rows):
        for k in range(rp[r], rp[r + 1]):
            m[r, ci[k]] += v[k]
    return m
