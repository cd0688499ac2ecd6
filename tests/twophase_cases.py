"""Shared helpers for the two-phase program (oracle and HIP runtime)."""
import ctypes as C

import numpy as np

from equelle_amd.capi import dptr, iptr
from equelle_amd.workloads import twophase_inputs, twophase_params


def oracle_residual(api, params, perm, poro, sc, sv, which, p, sw0, sw, dt):
    """(F, rp, ci, v) of the oracle's pressure (which=0) or transport (which=1) residual."""
    n = perm.size
    info = np.zeros(2, dtype=np.int32)
    args = (params, dptr(perm), dptr(poro), n, iptr(sc), dptr(sv), len(sc), which, dptr(p), dptr(sw0), dptr(sw), dt)
    api.twophase_residual_params(*args, iptr(info), None, None, None, None)
    F, rp = np.zeros(n), np.zeros(n + 1, dtype=np.int32)
    ci, v = np.zeros(info[1], dtype=np.int32), np.zeros(info[1])
    api.twophase_residual_params(*args, iptr(info), dptr(F), iptr(rp), iptr(ci), dptr(v))
    return F, rp, ci, v


def oracle_run(api, params, perm, poro, sw0, sc, sv, dts):
    n = perm.size
    p, sw = np.zeros(n), np.zeros(n)
    it = np.zeros(2 * len(dts), dtype=np.int32)
    api.twophase_run(params, dptr(perm), dptr(poro), dptr(sw0), n, iptr(sc), dptr(sv), len(sc),
                     dptr(np.asarray(dts, dtype=float)), len(dts), dptr(p), dptr(sw), iptr(it))
    return p, sw, it


def fi_params(nx, ny, nz, **kw):
    """Parameters of the fully implicit program: the workload's grid, viscosities and
    solver settings with the program's own linear relative permeabilities (corey=1,
    twophase_fully_implicit.equelle:62-78)."""
    return twophase_params(nx, ny, nz, **kw).replace("corey=2", "corey=1")


def oracle_fi_residual(api, params, perm, poro, sc, sv, p, sw0, sw, dt):
    """(F, rp, ci, v) of the oracle's fully implicit system residual at (p, sw): 2n rows
    in block order [pressure; transport], columns [p; sw]."""
    n = perm.size
    info = np.zeros(2, dtype=np.int32)
    args = (params, dptr(perm), dptr(poro), n, iptr(sc), dptr(sv), len(sc), dptr(p), dptr(sw0), dptr(sw), dt)
    api.twophase_fi_residual(*args, iptr(info), None, None, None, None)
    F, rp = np.zeros(2 * n), np.zeros(2 * n + 1, dtype=np.int32)
    ci, v = np.zeros(info[1], dtype=np.int32), np.zeros(info[1])
    api.twophase_fi_residual(*args, iptr(info), dptr(F), iptr(rp), iptr(ci), dptr(v))
    return F, rp, ci, v


def oracle_fi_run(api, params, perm, poro, sw0, sc, sv, dts):
    n = perm.size
    p, sw = np.zeros(n), np.zeros(n)
    it = np.zeros(len(dts), dtype=np.int32)
    api.twophase_fi_run(params, dptr(perm), dptr(poro), dptr(sw0), n, iptr(sc), dptr(sv), len(sc),
                        dptr(np.asarray(dts, dtype=float)), len(dts), dptr(p), dptr(sw), iptr(it))
    return p, sw, it


def case(nx, ny, nz, seed=42):
    perm, poro, sw0, sc, sv = twophase_inputs(nx, ny, nz, seed=seed)
    return perm, poro, sw0, sc, sv


__all__ = ["oracle_residual", "oracle_run", "oracle_fi_residual", "oracle_fi_run", "fi_params", "case",
           "twophase_params", "C"]
