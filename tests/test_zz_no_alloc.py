"""The allocation-free reduction (eqhip_reduce_noalloc) against eqhip_reduce.

Both run the same kernels in the same order (fixed grid for a given size), so every
result must be bit-identical.  Two calls in a row, the second on a smaller collection:
the fixed scratch keeps the first call's partials, and the second must not read them.
"""
import numpy as np
import pytest

from equelle_amd.capi import REDUCE, EquelleError, dptr, out_int

pytestmark = pytest.mark.gpu


def _reduce(api, fn, rt, h, kind):
    out = np.zeros(1)
    getattr(api, fn)(rt, REDUCE[kind], h, dptr(out))
    return float(out[0])


@pytest.mark.parametrize("sizes", [(3_000_017, 1_001), (200_000, 0), (65_536, 255)])
def test_noalloc_reduce_matches_reduce(hip_api, sizes):
    import ctypes as C
    rt = C.c_void_p()
    hip_api.runtime_create(b"grid_dim=2 nx=4 ny=3", C.byref(rt))
    try:
        rng = np.random.default_rng(sum(sizes))
        for n in sizes:
            print(f"no-alloc reduction: n={n}", flush=True)
            # values in [0.5, 1.5) keep the product finite for small n; large n overflow the
            # product to inf in both entry points alike
            x = rng.uniform(0.5, 1.5, n) * rng.choice([-1.0, 1.0], n)
            h = out_int()
            hip_api.scalar_from_host(rt, dptr(x) if n else None, n, C.byref(h))
            for kind in ("sum", "prod", "max", "min", "twonorm"):
                if n == 0 and kind in ("max", "min"):   # both entry points reject these
                    for fn in ("reduce", "reduce_noalloc"):
                        with pytest.raises(EquelleError, match="empty"):
                            _reduce(hip_api, fn, rt, h.value, kind)
                    continue
                a = _reduce(hip_api, "reduce", rt, h.value, kind)
                b = _reduce(hip_api, "reduce_noalloc", rt, h.value, kind)
                assert a == b or (np.isnan(a) and np.isnan(b)), (n, kind, a, b)
            if n:
                assert np.isclose(_reduce(hip_api, "reduce_noalloc", rt, h.value, "sum"), x.sum(), rtol=1e-12)
                assert _reduce(hip_api, "reduce_noalloc", rt, h.value, "max") == x.max()
                assert np.isclose(_reduce(hip_api, "reduce_noalloc", rt, h.value, "twonorm"),
                                  np.linalg.norm(x), rtol=1e-12)
            hip_api.release(rt, 0, h.value)
        h = out_int()
        x = np.ones(5)
        hip_api.scalar_from_host(rt, dptr(x), 5, C.byref(h))
        with pytest.raises(EquelleError, match="kind"):
            hip_api.reduce_noalloc(rt, 7, h.value, dptr(np.zeros(1)))
    finally:
        hip_api.runtime_destroy(rt)
