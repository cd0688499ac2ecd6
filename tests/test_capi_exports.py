"""CPU-side checks of the C-ABI library: it loads without a GPU and exports every
entry point include/eqhip.h declares (no compute calls are made here)."""
import os
import re

from conftest import HIP_SO, ROOT

from equelle_amd.capi import SIGNATURES, CApi


def declared():
    text = open(os.path.join(ROOT, "include", "eqhip.h")).read()
    return sorted(set(re.findall(r"\b(eqhip_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    import ctypes
    lib = ctypes.CDLL(HIP_SO)
    names = declared()
    assert len(names) >= 45
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_the_header():
    names = {n[len("eqhip_"):] for n in declared()} - {"last_error"}
    assert names <= set(SIGNATURES), names - set(SIGNATURES)
    api = CApi(HIP_SO, "eqhip_")
    assert not (set(api.missing) & names), set(api.missing) & names


def test_no_cpu_fallback_in_product_sources():
    """The product library never links or loads the oracle."""
    src = os.path.join(ROOT, "equelle_amd")
    for dirpath, _, files in os.walk(src):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".hpp", ".h", "Makefile")):
                text = open(os.path.join(dirpath, f), errors="ignore").read()
                assert "liboracle" not in text and "eqo_" not in text, f
