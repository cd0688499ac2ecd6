"""Python mirror of the ``equelleCUDA::EquelleRuntimeCUDA`` interface.

Method names, argument meaning and error behaviour follow the reference class
(backends/cuda/include/EquelleRuntimeCUDA.hpp:89-331); every call goes through
the C-ABI of the HIP runtime library (include/eqhip.h).  Values live on the
device: a :class:`CollOfScalar` is a handle to a device value array plus its
device CSR Jacobian.

Usage mirrors generated code (examples/app/out_heateq.cpp)::

    er = EquelleRuntimeCUDA("grid_dim=2 nx=100 ny=100 k=0.3 cv=1")
    ifaces = er.interiorFaces()
    itrans = k * (er.norm(ifaces) / er.norm(er.centroid(first) - er.centroid(second)))
    u = er.newtonSolve(lambda u: residual(u, u0, dt), u0)
"""
from __future__ import annotations

import ctypes as C
import os
import sys
import time
from typing import Callable, Optional, Union

import numpy as np

from .capi import (BUILTIN_SETS, CMPS, KIND_BOOL, KIND_SCALAR, KIND_SET, OPS,
                   REDUCE, CApi, EquelleError, dptr, iptr, uptr)

__all__ = ["EquelleRuntimeCUDA", "CollOfScalar", "CollOfIndices", "CollOfCell",
           "CollOfFace", "CollOfBool", "CollOfVector", "makeArray", "EquelleError",
           "load_library", "params_to_text"]

_LIB = None


def library_path() -> str:
    here = os.path.dirname(os.path.abspath(__file__))
    return os.path.join(here, "lib", "libeqhip.so")


def load_library() -> CApi:
    """Load the HIP runtime library (fails loudly when it has not been built)."""
    global _LIB
    if _LIB is None:
        path = library_path()
        if not os.path.exists(path):
            raise EquelleError(
                f"HIP runtime library missing at {path}; run __graft_entry__.build() first")
        _LIB = CApi(path, "eqhip_")
    return _LIB


def params_to_text(params) -> bytes:
    if isinstance(params, (bytes, str)):
        return params.encode() if isinstance(params, str) else params
    return " ".join(f"{k}={v}" for k, v in params.items()).encode()


class _Handle:
    __slots__ = ("er", "h", "kind")

    def __init__(self, er: "EquelleRuntimeCUDA", h: int, kind: int):
        self.er, self.h, self.kind = er, h, kind

    def __del__(self):
        er = getattr(self, "er", None)
        if er is not None and er._rt is not None and sys is not None:
            try:
                er.api.release(er._rt, self.kind, self.h)
            except Exception:
                pass


class CollOfBool(_Handle):
    """Device collection of booleans (``thrust::device_vector<bool>`` in the reference)."""

    def __init__(self, er, h):
        super().__init__(er, h, KIND_BOOL)

    def size(self) -> int:
        n = C.c_int()
        self.er.api.bool_size(self.er._rt, self.h, C.byref(n))
        return n.value

    def __len__(self):
        return self.size()

    def to_numpy(self) -> np.ndarray:
        out = np.zeros(self.size(), dtype=np.uint8)
        self.er.api.bool_copy(self.er._rt, self.h, uptr(out))
        return out.astype(bool)


class CollOfIndices(_Handle):
    """Index set on cells (codim 0) or faces (codim 1); ``-1`` marks an empty entity."""

    def __init__(self, er, h, codim: int, full: bool = False):
        super().__init__(er, h, KIND_SET)
        self.codim = codim
        self.full = full

    def isFull(self) -> bool:
        return self.full

    def size(self) -> int:
        n = C.c_int()
        self.er.api.set_size(self.er._rt, self.h, C.byref(n))
        return n.value

    def __len__(self):
        return self.size()

    def stdToHost(self) -> np.ndarray:
        out = np.zeros(self.size(), dtype=np.int32)
        if out.size:
            self.er.api.set_copy(self.er._rt, self.h, iptr(out))
        return out


CollOfCell = CollOfIndices
CollOfFace = CollOfIndices


class CollOfScalar(_Handle):
    """Values plus (optionally) a CSR Jacobian with respect to the primary variable."""

    def __init__(self, er, h):
        super().__init__(er, h, KIND_SCALAR)

    # ---- introspection ----
    def _info(self):
        info = np.zeros(5, dtype=np.int32)
        self.er.api.scalar_info(self.er._rt, self.h, iptr(info))
        return info

    def size(self) -> int:
        return int(self._info()[0])

    def __len__(self):
        return self.size()

    def useAutoDiff(self) -> bool:
        return bool(self._info()[1])

    def value(self) -> np.ndarray:
        out = np.zeros(self.size(), dtype=np.float64)
        if out.size:
            self.er.api.scalar_copy(self.er._rt, self.h, dptr(out))
        return out

    copyToHost = value

    def derivative(self):
        """Host copy of the Jacobian as ``(rowPtr, colInd, vals, rows, cols)``;
        ``None`` for an empty (constant) Jacobian."""
        info = self._info()
        rows, cols, nnz = int(info[2]), int(info[3]), int(info[4])
        if rows == 0:
            return None
        rp = np.zeros(rows + 1, dtype=np.int32)
        ci = np.zeros(max(nnz, 1), dtype=np.int32)
        v = np.zeros(max(nnz, 1), dtype=np.float64)
        self.er.api.jacobian_copy(self.er._rt, self.h, iptr(rp), iptr(ci), dptr(v))
        return rp, ci[:nnz], v[:nnz], rows, cols

    matrixToHost = derivative

    # ---- arithmetic (CollOfScalar.hpp:187-412) ----
    def _bin(self, op, other, reverse=False):
        er = self.er
        out = C.c_int()
        if isinstance(other, CollOfScalar):
            a, b = (other, self) if reverse else (self, other)
            er.api.binop(er._rt, OPS[op], a.h, b.h, C.byref(out))
            return CollOfScalar(er, out.value)
        s = float(other)
        if op == "*":
            er.api.scalar_binop(er._rt, 2, s, self.h, 1, C.byref(out))
        elif op == "/":
            er.api.scalar_binop(er._rt, 3, s, self.h, 1 if reverse else 0, C.byref(out))
        elif op == "+" or op == "-":
            # Scalar +/- collection is not part of the reference API; generated code
            # writes (s Extend set) instead.  Keep the same restriction.
            raise TypeError("Scalar +/- CollOfScalar is not an Equelle operation; use operatorExtend")
        return CollOfScalar(er, out.value)

    def __add__(self, o): return self._bin("+", o)
    def __sub__(self, o): return self._bin("-", o)
    def __mul__(self, o): return self._bin("*", o)
    def __rmul__(self, o): return self._bin("*", o, reverse=True)
    def __truediv__(self, o): return self._bin("/", o)
    def __rtruediv__(self, o): return self._bin("/", o, reverse=True)

    def __neg__(self):
        out = C.c_int()
        self.er.api.negate(self.er._rt, self.h, C.byref(out))
        return CollOfScalar(self.er, out.value)

    def _cmp(self, op, other, swapped=False):
        er = self.er
        out = C.c_int()
        if isinstance(other, CollOfScalar):
            er.api.compare(er._rt, CMPS[op], self.h, other.h, C.byref(out))
        else:
            er.api.compare_scalar(er._rt, CMPS[op], self.h, float(other), 1 if swapped else 0, C.byref(out))
        return CollOfBool(er, out.value)

    def __gt__(self, o): return self._cmp(">", o)
    def __ge__(self, o): return self._cmp(">=", o)
    def __lt__(self, o): return self._cmp("<", o)
    def __le__(self, o): return self._cmp("<=", o)
    def __eq__(self, o): return self._cmp("==", o)
    def __ne__(self, o): return self._cmp("!=", o)
    __hash__ = None


class CollOfVector:
    """Collection of `dim`-vectors (setup-time geometry, no AD), host resident,
    like the reference's CollOfVector (CollOfVector.hpp:38-168)."""

    def __init__(self, er, data: np.ndarray):
        self.er = er
        self.data = np.asarray(data, dtype=np.float64)

    def dim(self):
        return self.data.shape[1]

    def numVectors(self):
        return self.data.shape[0]

    def __getitem__(self, d) -> CollOfScalar:
        return self.er._scalar(np.ascontiguousarray(self.data[:, d]))

    col = __getitem__

    def __sub__(self, o): return CollOfVector(self.er, self.data - o.data)
    def __add__(self, o): return CollOfVector(self.er, self.data + o.data)
    def __neg__(self): return CollOfVector(self.er, -self.data)

    def __mul__(self, s):
        if isinstance(s, CollOfScalar):
            return CollOfVector(self.er, self.data * s.value()[:, None])
        return CollOfVector(self.er, self.data * float(s))

    __rmul__ = __mul__

    def __truediv__(self, s):
        if isinstance(s, CollOfScalar):
            return CollOfVector(self.er, self.data / s.value()[:, None])
        return CollOfVector(self.er, self.data / float(s))

    def norm(self) -> CollOfScalar:
        return self.er._scalar(np.sqrt((self.data * self.data).sum(axis=1)))

    def dot(self, o) -> CollOfScalar:
        return self.er._scalar((self.data * o.data).sum(axis=1))


def makeArray(*colls):
    """EquelleRuntimeCUDA.hpp:41-72: arrays of collections become tuples."""
    return tuple(colls)


class NewtonStats:
    def __init__(self):
        self.iterations = 0
        self.initial_norm = 0.0
        self.final_norm = 0.0
        self.linear_iterations = 0
        self.stop_reason = 0
        self.seconds = 0.0


class EquelleRuntimeCUDA:
    """Host mirror of the reference runtime class over the C-ABI."""

    def __init__(self, params="", api: Optional[CApi] = None, param_files=()):
        self.api = api if api is not None else load_library()
        self.params = {}
        for path in param_files:
            self.params.update(_read_param_file(path))
        if isinstance(params, dict):
            self.params.update({k: str(v) for k, v in params.items()})
        else:
            for tok in (params.decode() if isinstance(params, bytes) else params).split():
                if "=" in tok:
                    k, v = tok.split("=", 1)
                    self.params[k] = v
        rt = C.c_void_p()
        self.api.runtime_create(params_to_text(self.params), C.byref(rt))
        self._rt = rt
        info = np.zeros(5, dtype=np.int32)
        self.api.grid_info(self._rt, iptr(info))
        self.dimensions, self.nc, self.nf, self.nif, self.nbf = (int(x) for x in info)
        self.verbose = int(self.params.get("verbose", 0))
        self.max_iter = int(self.params.get("max_iter", 10))
        self.abs_res_tol = float(self.params.get("abs_res_tol", 1e-6))
        self.rel_res_tol = float(self.params.get("rel_res_tol", 0.0))
        self.output_to_file = self.params.get("output_to_file", "false") in ("1", "true")
        self._outputcount = {}
        self.last_newton = NewtonStats()

    def close(self):
        if self._rt is not None:
            self.api.runtime_destroy(self._rt)
            self._rt = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------- helpers ----------------
    def _scalar(self, values) -> CollOfScalar:
        v = np.ascontiguousarray(values, dtype=np.float64)
        out = C.c_int()
        self.api.scalar_from_host(self._rt, dptr(v), int(v.size), C.byref(out))
        return CollOfScalar(self, out.value)

    def collOfScalar(self, values) -> CollOfScalar:
        """CollOfScalar(const std::vector<double>&)."""
        return self._scalar(values)

    def collOfScalarWithJacobian(self, values, rp, ci, v, cols) -> CollOfScalar:
        """CollOfScalar(CudaArray, CudaMatrix) from host CSR arrays (test hook)."""
        vals = np.ascontiguousarray(values, dtype=np.float64)
        rp = np.ascontiguousarray(rp, dtype=np.int32)
        ci = np.ascontiguousarray(ci, dtype=np.int32)
        v = np.ascontiguousarray(v, dtype=np.float64)
        out = C.c_int()
        self.api.scalar_with_jacobian(self._rt, dptr(vals), int(vals.size), iptr(rp), iptr(ci), dptr(v),
                                      int(cols), C.byref(out))
        return CollOfScalar(self, out.value)

    def primaryVariable(self, x: CollOfScalar) -> CollOfScalar:
        """CollOfScalar(const CollOfScalar&, bool primaryVariable=true)."""
        out = C.c_int()
        self.api.primary(self._rt, x.h, C.byref(out))
        return CollOfScalar(self, out.value)

    def _set(self, name):
        out = C.c_int()
        self.api.set_builtin(self._rt, BUILTIN_SETS[name], C.byref(out))
        codim = 0 if "Cells" in name else 1
        return CollOfIndices(self, out.value, codim, full=name in ("allCells", "allFaces"))

    def indices(self, idx, codim) -> CollOfIndices:
        a = np.ascontiguousarray(idx, dtype=np.int32)
        out = C.c_int()
        self.api.set_from_host(self._rt, codim, iptr(a), int(a.size), C.byref(out))
        return CollOfIndices(self, out.value, codim)

    # ---------------- topology (EquelleRuntimeCUDA.hpp:98-122) ----------------
    def allCells(self): return self._set("allCells")
    def boundaryCells(self): return self._set("boundaryCells")
    def interiorCells(self): return self._set("interiorCells")
    def allFaces(self): return self._set("allFaces")
    def boundaryFaces(self): return self._set("boundaryFaces")
    def interiorFaces(self): return self._set("interiorFaces")

    def firstCell(self, faces: CollOfIndices) -> CollOfIndices:
        out = C.c_int()
        self.api.first_cell(self._rt, faces.h, C.byref(out))
        return CollOfIndices(self, out.value, 0)

    def secondCell(self, faces: CollOfIndices) -> CollOfIndices:
        out = C.c_int()
        self.api.second_cell(self._rt, faces.h, C.byref(out))
        return CollOfIndices(self, out.value, 0)

    def isEmpty(self, s: CollOfIndices) -> CollOfBool:
        out = C.c_int()
        self.api.is_empty(self._rt, s.h, C.byref(out))
        return CollOfBool(self, out.value)

    # ---------------- geometry ----------------
    def norm(self, x):
        """|set| (volumes/areas), the 2-norm of a CollOfVector, or the elementwise absolute
        value of a CollOfScalar (EquelleRuntimeCPU.cpp:241-254)."""
        if isinstance(x, CollOfVector):
            return x.norm()
        out = C.c_int()
        if isinstance(x, CollOfScalar):
            self.api.norm_scalar(self._rt, x.h, C.byref(out))
            return CollOfScalar(self, out.value)
        self.api.norm_set(self._rt, x.h, C.byref(out))
        return CollOfScalar(self, out.value)

    def centroid(self, s: CollOfIndices) -> CollOfVector:
        out = np.zeros((s.size(), self.dimensions), dtype=np.float64)
        if out.size:
            self.api.centroid(self._rt, s.h, dptr(out))
        return CollOfVector(self, out)

    def normal(self, faces: CollOfIndices) -> CollOfVector:
        out = np.zeros((faces.size(), self.dimensions), dtype=np.float64)
        if out.size:
            self.api.normal(self._rt, faces.h, dptr(out))
        return CollOfVector(self, out)

    def dot(self, v1: CollOfVector, v2: CollOfVector) -> CollOfScalar:
        return v1.dot(v2)

    # ---------------- operators ----------------
    def gradient(self, x: CollOfScalar) -> CollOfScalar:
        out = C.c_int()
        self.api.gradient(self._rt, x.h, C.byref(out))
        return CollOfScalar(self, out.value)

    def divergence(self, x: CollOfScalar) -> CollOfScalar:
        out = C.c_int()
        self.api.divergence(self._rt, x.h, C.byref(out))
        return CollOfScalar(self, out.value)

    def negGradient(self, x: CollOfScalar) -> CollOfScalar:
        """ngrad * x (EquelleRuntimeCPU.cpp:345-348)."""
        out = C.c_int()
        self.api.neg_gradient(self._rt, x.h, C.byref(out))
        return CollOfScalar(self, out.value)

    def interiorDivergence(self, x: CollOfScalar) -> CollOfScalar:
        """div * x on InteriorFaces() (EquelleRuntimeCPU.cpp:362-365)."""
        out = C.c_int()
        self.api.interior_divergence(self._rt, x.h, C.byref(out))
        return CollOfScalar(self, out.value)

    def operatorExtend(self, data, a, b=None):
        out = C.c_int()
        if b is None:   # operatorExtend(Scalar, set)
            self.api.extend_scalar(self._rt, float(data), a.h, C.byref(out))
        else:
            self.api.extend(self._rt, data.h, a.h, b.h, C.byref(out))
        return CollOfScalar(self, out.value)

    def operatorOn(self, data, from_set: CollOfIndices, to_set: CollOfIndices):
        out = C.c_int()
        if isinstance(data, CollOfIndices):
            self.api.set_on(self._rt, data.h, from_set.h, to_set.h, C.byref(out))
            return CollOfIndices(self, out.value, data.codim)
        self.api.on(self._rt, data.h, from_set.h, to_set.h, C.byref(out))
        return CollOfScalar(self, out.value)

    def trinaryIf(self, pred: CollOfBool, t, f):
        out = C.c_int()
        if isinstance(t, CollOfIndices):
            self.api.set_trinary(self._rt, pred.h, t.h, f.h, C.byref(out))
            return CollOfIndices(self, out.value, t.codim)
        self.api.trinary(self._rt, pred.h, t.h, f.h, C.byref(out))
        return CollOfScalar(self, out.value)

    def sqrt(self, x: CollOfScalar) -> CollOfScalar:
        out = C.c_int()
        self.api.sqrt(self._rt, x.h, C.byref(out))
        return CollOfScalar(self, out.value)

    # ---------------- reductions ----------------
    def _reduce(self, kind, x):
        out = C.c_double()
        self.api.reduce(self._rt, REDUCE[kind], x.h, C.byref(out))
        return out.value

    def minReduce(self, x): return self._reduce("min", x)
    def maxReduce(self, x): return self._reduce("max", x)
    def sumReduce(self, x): return self._reduce("sum", x)
    def prodReduce(self, x): return self._reduce("prod", x)
    def twoNorm(self, x): return self._reduce("twonorm", x)
    twoNormTester = twoNorm

    # ---------------- solver (EquelleRuntimeCUDA_impl.hpp:216-300) ----------------
    def solveForUpdate(self, residual: CollOfScalar):
        du, it = C.c_int(), C.c_int()
        self.api.solve_for_update(self._rt, residual.h, C.byref(du), C.byref(it))
        return CollOfScalar(self, du.value), it.value

    def newtonSolve(self, rescomp: Callable[[CollOfScalar], CollOfScalar], u_initialguess: CollOfScalar):
        t0 = time.perf_counter()
        st = NewtonStats()
        u = self.primaryVariable(u_initialguess)
        residual = rescomp(u)
        r0 = self.twoNorm(residual)
        rn = r0

        def done(r):
            if r <= self.abs_res_tol:
                return 1
            if self.rel_res_tol > 0.0 and r <= self.rel_res_tol * r0:
                return 2
            return 0

        it = 0
        while not done(rn) and it < self.max_iter:
            du, lit = self.solveForUpdate(residual)
            st.linear_iterations += lit
            u = u - du
            residual = rescomp(u)
            rn = self.twoNorm(residual)
            it += 1
            if self.verbose > 1:
                print(f"    newtonSolve: iter = {it} (max = {self.max_iter}), norm(residual) = {rn} "
                      f"(tol = {self.abs_res_tol})")
        st.iterations, st.initial_norm, st.final_norm, st.stop_reason = it, r0, rn, done(rn)
        st.seconds = time.perf_counter() - t0
        if self.verbose > 0:
            if not st.stop_reason:
                print(f"Newton solver failed to converge in {self.max_iter} iterations")
            else:
                print(f"Newton solver converged in {it} iterations")
        self.last_newton = st
        return self._scalar(u.value())

    # ---------------- input / output ----------------
    def inputScalarWithDefault(self, name, default):
        return float(self.params.get(name, default))

    def inputCollectionOfScalar(self, name, coll: CollOfIndices) -> CollOfScalar:
        size = coll.size()
        if self.params.get(name + "_from_file", "false") in ("1", "true"):
            fname = self.params.get(name + "_filename")
            if fname is None or not os.path.exists(fname):
                raise EquelleError(f"Could not find file {fname}")
            data = np.array(open(fname).read().split(), dtype=np.float64)
            if data.size != size:
                raise EquelleError(f"Unexpected size of input data for {name} in file {fname}")
            return self._scalar(data)
        if name not in self.params:
            raise EquelleError(f"parameter {name} not found")
        return self._scalar(np.full(size, float(self.params[name])))

    def inputDomainSubsetOf(self, name, superset: CollOfIndices) -> CollOfIndices:
        fname = self.params.get(name + "_filename")
        if fname is None or not os.path.exists(fname):
            raise EquelleError(f"Could not find file {fname}")
        data = np.array(open(fname).read().split(), dtype=np.int64).astype(np.int32)
        if np.any(np.diff(data) < 0):
            raise EquelleError(f"Input set {name} was not sorted in ascending order")
        sub = self.indices(data, superset.codim)
        self.api.set_contains(self._rt, superset.h, sub.h)
        return sub

    def inputSequenceOfScalar(self, name):
        fname = self.params.get(name + "_filename")
        if fname is None or not os.path.exists(fname):
            raise EquelleError(f"Could not find file {fname}")
        return [float(x) for x in open(fname).read().split()]

    def output(self, tag, x):
        if not isinstance(x, CollOfScalar):
            print(f"{tag} = {x}")
            return
        vals = x.value()
        if self.output_to_file:
            count = self._outputcount.get(tag, 0)
            self._outputcount[tag] = count + 1
            with open(f"{tag}-{count:05d}.output", "w") as f:
                for v in vals:
                    f.write(f"{v:.16g}\n")
        else:
            print(f"\nValues in {tag}\n" + "  ".join(f"{v:g}" for v in vals))

    def ensureGridDimensionMin(self, d):
        if self.dimensions < d:
            raise EquelleError(f"Equelle simulator requires minimum {d} dimensions, "
                               f"but grid only has {self.dimensions} dimensions.")

    # ---------------- test hooks (EquelleRuntimeCUDA.hpp:296-305) ----------------
    def _opmat(self, which):
        info = np.zeros(3, dtype=np.int32)
        self.api.opmat_info(self._rt, which, iptr(info))
        rows, cols, nnz = (int(x) for x in info)
        rp = np.zeros(rows + 1, dtype=np.int32)
        ci = np.zeros(nnz, dtype=np.int32)
        v = np.zeros(nnz, dtype=np.float64)
        self.api.opmat_copy(self._rt, which, iptr(rp), iptr(ci), dptr(v))
        return rp, ci, v, rows, cols

    def getGradMatrix(self): return self._opmat(0)
    def getDivMatrix(self): return self._opmat(1)
    def getFulldivMatrix(self): return self._opmat(2)

    def gridIntArray(self, which):
        n = {0: 2 * self.nf, 1: self.nc + 1, 2: 2 * self.dimensions * self.nc}[which]
        out = np.zeros(n, dtype=np.int32)
        self.api.grid_int_array(self._rt, which, iptr(out))
        return out

    def gridDoubleArray(self, which):
        d = self.dimensions
        n = {0: d * self.nc, 1: self.nc, 2: d * self.nf, 3: self.nf, 4: d * self.nf}[which]
        out = np.zeros(n, dtype=np.float64)
        self.api.grid_double_array(self._rt, which, dptr(out))
        return out


def _read_param_file(path):
    """Opm::ParameterGroup .param file: one key=value per line, '#' comments."""
    out = {}
    with open(path) as f:
        for line in f:
            line = line.split("#", 1)[0].strip()
            if "=" in line:
                k, v = line.split("=", 1)
                out[k.strip()] = v.strip()
    return out
