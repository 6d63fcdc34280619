"""Shared test setup.

* registers the ``gpu`` marker (tests that need an MI355X);
* puts the package directory and the repo root on sys.path;
* provides the CPU oracle (built from oracle/ if it is missing) — the oracle is
  only ever the checker in these tests.
"""
import ctypes as C
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "hpr-socp_amd")
for p in (PKG, ROOT, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


def _oracle_lib():
    path = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(path):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True,
                       stdout=subprocess.DEVNULL)
    from hprsocp import _abi
    return _abi.bind(C.CDLL(path), "oracle_")


@pytest.fixture(scope="session")
def oracle_lib():
    return _oracle_lib()


@pytest.fixture(scope="session")
def gpu_lib():
    """The HIP backend; builds it in-tree when absent (hipcc ships with ROCm)."""
    path = os.path.join(PKG, "hprsocp", "libhprsocp.so")
    if not os.path.exists(path):
        subprocess.run(["make", "-C", os.path.join(PKG, "csrc"), "-j8"], check=True,
                       stdout=subprocess.DEVNULL)
    import hprsocp
    return hprsocp.load_library()
