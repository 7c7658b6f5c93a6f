"""Shared fixtures: the HIP engine and the two CPU oracles behind the same
C-ABI (include/gmes_engine.h)."""

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

ORACLE_LIB = os.path.join(ROOT, "oracle", "libgmes_oracle.so")
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libgmes_ref.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


@pytest.fixture(scope="session")
def oracle_lib():
    from gmes_amd._abi import Library
    if not os.path.exists(ORACLE_LIB):
        import subprocess
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "oracle"], check=True)
    return Library(ORACLE_LIB)


@pytest.fixture(scope="session")
def ref_lib():
    from gmes_amd._abi import Library
    if not os.path.exists(REF_LIB):
        pytest.skip("reference harness oracle/_ref/libgmes_ref.so not built")
    return Library(REF_LIB)


@pytest.fixture(scope="session")
def engine_lib():
    from gmes_amd import _abi
    return _abi.engine()
