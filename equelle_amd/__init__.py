"""equelle_amd — an MI355X-native (gfx950) runtime for Equelle simulators.

The product is the C++/HIP library ``equelle_amd/lib/libeqhip.so`` (built from
``equelle_amd/csrc``) with its C++ facade ``include/EquelleRuntimeCUDA.hpp`` and
C-ABI ``include/eqhip.h``.  This package holds the Python mirror of the runtime
interface (:mod:`equelle_amd.runtime`) and the workload generators
(:mod:`equelle_amd.workloads`).
"""
from .capi import EquelleError  # noqa: F401
