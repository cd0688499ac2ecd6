"""Shared fixtures.

* ``oracle_api``  — the CPU oracle (oracle/liboracle.so) behind the same C-ABI
  shape as the HIP runtime, prefix ``eqo_``.  Only tests use it.
* ``hip_api``     — the HIP runtime (equelle_amd/lib/libeqhip.so), prefix ``eqhip_``.
  GPU tests are marked ``@pytest.mark.gpu``.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from equelle_amd.capi import CApi  # noqa: E402

ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
HIP_SO = os.path.join(ROOT, "equelle_amd", "lib", "libeqhip.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP runtime C-ABI)")


@pytest.fixture(scope="session")
def oracle_api():
    if not os.path.exists(ORACLE_SO):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    return CApi(ORACLE_SO, "eqo_")


@pytest.fixture(scope="session")
def hip_api():
    if not os.path.exists(HIP_SO):
        pytest.fail(f"HIP runtime library missing: {HIP_SO} (run __graft_entry__.build())")
    return CApi(HIP_SO, "eqhip_")
