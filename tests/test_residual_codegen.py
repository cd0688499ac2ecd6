"""The fused residual kernel's lowering builds for gfx950 without a device: the run-time
compiler takes the generated source of a heat-like residual program (gradient,
divergence, On / Extend, elementwise AD), with and without derivatives."""
import ctypes as C

from equelle_amd.runtime import load_library


def test_residual_kernel_builds_for_gfx950():
    lib = load_library()
    log = C.create_string_buffer(1 << 16)
    assert lib.residual_codegen_selftest(log, len(log)) == 0, log.value.decode(errors="replace")
