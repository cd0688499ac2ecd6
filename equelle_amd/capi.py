"""ctypes binding of the handle-based Equelle runtime C-ABI (include/eqhip.h).

The binding is generic over the symbol prefix: the product library exports the
entry points as ``eqhip_<name>``.  Every call returns 0 on success; a nonzero
status raises :class:`EquelleError` carrying ``<prefix>last_error()``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_P = C.c_void_p
_I = C.c_int
_D = C.c_double
_IP = C.POINTER(C.c_int)
_DP = C.POINTER(C.c_double)
_UP = C.POINTER(C.c_ubyte)
_S = C.c_char_p

# name -> argtypes (restype is int unless listed in _RESTYPES)
SIGNATURES = {
    "runtime_create": [_S, C.POINTER(_P)],
    "runtime_destroy": [_P],
    "grid_info": [_P, _IP],
    "grid_int_array": [_P, _I, _IP],
    "grid_double_array": [_P, _I, _DP],
    "opmat_info": [_P, _I, _IP],
    "opmat_copy": [_P, _I, _IP, _IP, _DP],
    "set_builtin": [_P, _I, _IP],
    "set_from_host": [_P, _I, _IP, _I, _IP],
    "set_size": [_P, _I, _IP],
    "set_copy": [_P, _I, _IP],
    "first_cell": [_P, _I, _IP],
    "second_cell": [_P, _I, _IP],
    "set_on": [_P, _I, _I, _I, _IP],
    "set_trinary": [_P, _I, _I, _I, _IP],
    "is_empty": [_P, _I, _IP],
    "set_contains": [_P, _I, _I],
    "bool_size": [_P, _I, _IP],
    "bool_copy": [_P, _I, _UP],
    "compare": [_P, _I, _I, _I, _IP],
    "compare_scalar": [_P, _I, _I, _D, _I, _IP],
    "scalar_from_host": [_P, _DP, _I, _IP],
    "scalar_uniform": [_P, _I, _D, _IP],
    "scalar_with_jacobian": [_P, _DP, _I, _IP, _IP, _DP, _I, _IP],
    "primary": [_P, _I, _IP],
    "scalar_info": [_P, _I, _IP],
    "scalar_copy": [_P, _I, _DP],
    "jacobian_copy": [_P, _I, _IP, _IP, _DP],
    "release": [_P, _I, _I],
    "binop": [_P, _I, _I, _I, _IP],
    "scalar_binop": [_P, _I, _D, _I, _I, _IP],
    "negate": [_P, _I, _IP],
    "sqrt": [_P, _I, _IP],
    "trinary": [_P, _I, _I, _I, _IP],
    "gradient": [_P, _I, _IP],
    "divergence": [_P, _I, _IP],
    "neg_gradient": [_P, _I, _IP],
    "interior_divergence": [_P, _I, _IP],
    "norm_scalar": [_P, _I, _IP],
    "extend": [_P, _I, _I, _I, _IP],
    "extend_scalar": [_P, _D, _I, _IP],
    "on": [_P, _I, _I, _I, _IP],
    "reduce": [_P, _I, _I, _DP],
    "reduce_noalloc": [_P, _I, _I, _DP],
    "norm_set": [_P, _I, _IP],
    "centroid": [_P, _I, _DP],
    "normal": [_P, _I, _DP],
    "solve_for_update": [_P, _I, _IP, _IP],
    "csr_binop": [_I, _I, _I, _IP, _IP, _DP, _I, _I, _IP, _IP, _DP, _IP, _IP, _IP, _DP, _I],
    "csr_spmv": [_I, _I, _IP, _IP, _DP, _DP, _DP, _I],
    "sell_spmv": [_P, _I, _DP, _DP, _I],
    "heat_create": [_S, _DP, _I, _IP, _DP, _I, C.POINTER(_P)],
    "heat_step": [_P, _D, _IP, _DP],
    "heat_state": [_P, _DP],
    "heat_set_state": [_P, _DP],
    "heat_eval_time": [_P, _D, _I, _DP, _IP],
    "heat_destroy": [_P],
    "residual_stats": [C.POINTER(C.c_long), _DP],
    "residual_codegen_selftest": [C.c_char_p, _I],
    "batch_create": [_S, C.POINTER(_P)],
    "batch_destroy": [_P],
    "batch_add": [_P, _I, _I, _D, _D, _DP, _IP],
    "batch_remove": [_P, _I],
    "batch_begin_step": [_P, _I, _D],
    "batch_add_twophase": [_P, _S, _I, _I, _I, _DP, _DP, _DP, _I, _IP, _DP, _IP],
    "batch_twophase_state": [_P, _I, _DP, _DP],
    "batch_twophase_stats": [_P, _I, _IP],
    "batch_twophase_residual": [_P, _I, _I, _DP, _DP, _DP, _D, _IP, _DP, _IP, _IP, _DP],
    "batch_iterate": [_P, _IP, _I, _IP],
    "batch_in_flight": [_P, _I, _IP],
    "batch_step": [_P, _IP, _I, _D],
    "batch_stats": [_P, _I, _IP, _DP],
    "batch_stats_many": [_P, _IP, _I, _IP],
    "batch_state": [_P, _I, _DP],
    "batch_set_state": [_P, _I, _DP],
    "batch_set_twophase_state": [_P, _I, _DP, _DP],
    "batch_bytes": [_P, _I, _DP],
    "batch_kernel_time": [_P, _S, _I, _DP, _DP],
    "batch_shape": [_P, _IP],
    "batch_shape_classes": [_P, _IP],
    "batch_residual": [_P, _I, _DP, _DP, _D, _IP, _DP, _IP, _IP, _DP],
    "twophase_create": [_S, _DP, _DP, _DP, _I, _IP, _DP, _I, C.POINTER(_P)],
    "twophase_step": [_P, _D, _IP, _DP],
    "twophase_state": [_P, _DP, _DP],
    "twophase_set_state": [_P, _DP, _DP],
    "twophase_residual": [_P, _I, _DP, _DP, _DP, _D, _IP, _DP, _IP, _IP, _DP],
    "twophase_destroy": [_P],
    "server_create": [_S, C.POINTER(_P)],
    "server_destroy": [_P],
    "server_submit": [_P, _I, _I, _D, _D, _DP, _I, _D, _IP],
    "server_progress": [_P, _IP, _I, _IP],
    "server_submit_twophase": [_P, _S, _I, _I, _I, _DP, _DP, _DP, _I, _IP, _DP, _I, _D, _IP],
    "server_result": [_P, _I, _DP, _DP, _IP, _DP],
    "server_status": [_P, _IP, _DP],
    "server_set_max_newton": [_P, _I, _I],
    "timing": [_I],
    "timing_query": [_S, _IP, _DP],
    "device_sync": [],
    "heat_newton_sample": [_S, _DP, _I, _I, _D, _DP, _IP],
    "heat_run": [_S, _DP, _I, _IP, _DP, _I, _DP, _I, _DP, _IP, _DP],
    "twophase_run": [_S, _DP, _DP, _DP, _I, _IP, _DP, _I, _DP, _I, _DP, _DP, _IP],
    "twophase_residual_params": [_S, _DP, _DP, _I, _IP, _DP, _I, _I, _DP, _DP, _DP, _D, _IP, _DP, _IP, _IP, _DP],
    "twophase_fi_run": [_S, _DP, _DP, _DP, _I, _IP, _DP, _I, _DP, _I, _DP, _DP, _IP],
    "twophase_fi_residual": [_S, _DP, _DP, _I, _IP, _DP, _I, _DP, _DP, _DP, _D, _IP, _DP, _IP, _IP, _DP],
    "heat_residual": [_S, _DP, _DP, _I, _IP, _DP, _I, _D, _IP, _DP, _IP, _IP, _DP],
    "heat_residual_values": [_S, _DP, _DP, _I, _IP, _DP, _I, _D, _DP],
}
_RESTYPES = {"last_error": _S}

OPS = {"+": 0, "-": 1, "*": 2, "/": 3}
CMPS = {">": 0, ">=": 1, "<": 2, "<=": 3, "==": 4, "!=": 5}
REDUCE = {"sum": 0, "prod": 1, "max": 2, "min": 3, "twonorm": 4}
BUILTIN_SETS = {"allCells": 0, "boundaryCells": 1, "interiorCells": 2,
                "allFaces": 3, "boundaryFaces": 4, "interiorFaces": 5}
KIND_SCALAR, KIND_SET, KIND_BOOL = 0, 1, 2


class EquelleError(RuntimeError):
    """An error reported through the C-ABI (the reference throws std::runtime_error)."""


class CApi:
    """Resolved entry points of one library under one prefix."""

    def __init__(self, path: str, prefix: str, optional=()):
        if not os.path.exists(path):
            raise FileNotFoundError(f"Equelle runtime library not built: {path}")
        self.path = path
        self.lib = C.CDLL(path)
        self.prefix = prefix
        self.missing = []
        fn = getattr(self.lib, prefix + "last_error")
        fn.restype = _S
        fn.argtypes = []
        self._last_error = fn
        for name, args in SIGNATURES.items():
            try:
                f = getattr(self.lib, prefix + name)
            except AttributeError:
                self.missing.append(name)
                continue
            f.argtypes = args
            f.restype = _I
            setattr(self, name, self._checked(f, name))

    def _checked(self, f, name):
        def call(*args):
            rc = f(*args)
            if rc != 0:
                msg = self._last_error()
                raise EquelleError(f"{self.prefix}{name}: {msg.decode() if msg else 'error'}")
            return rc
        call.__name__ = name
        return call


def iptr(a: np.ndarray):
    return a.ctypes.data_as(_IP)


def dptr(a: np.ndarray):
    return a.ctypes.data_as(_DP)


def uptr(a: np.ndarray):
    return a.ctypes.data_as(_UP)


def out_int():
    return C.c_int(0)
