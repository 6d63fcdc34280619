"""hprsocp_check_restart / hprsocp_init_restart of the HIP library (host code,
no device needed) against the oracle's check_restart (algorithm.jl:1016-1099)
on random gap/KKT sequences: every field identical after every call."""
import ctypes as C
import os

import numpy as np
import pytest

from hprsocp import _abi


@pytest.fixture(scope="module")
def hip_host_lib():
    path = _abi.library_path("libhprsocp.so")
    if not os.path.exists(path):
        pytest.skip("libhprsocp.so not built (run __graft_entry__.build())")
    return _abi.bind(C.CDLL(path), "hprsocp_")


@pytest.mark.parametrize("seed", range(6))
def test_check_restart_sequences(hip_host_lib, oracle_lib, seed):
    rng = np.random.default_rng(seed)
    a, b = _abi.RestartInfo(), _abi.RestartInfo()
    hip_host_lib.hprsocp_init_restart(C.byref(a))
    oracle_lib.oracle_init_restart(C.byref(b))
    assert a.as_dict() == b.as_dict()
    check_iter = int(rng.choice([50, 150]))
    gap = 1.0
    flags = set()
    for it in range(0, 60 * check_iter, 25):
        gap *= float(np.exp(rng.normal(-0.05, 0.4)))
        cur = -gap if rng.random() < 0.02 else gap
        for r in (a, b):
            r.current_gap = cur
            r.inner = int(r.inner) + 25
        kkt = float(np.exp(rng.normal(-3, 1)))
        sigma = float(np.exp(rng.normal(0, 1)))
        hip_host_lib.hprsocp_check_restart(C.byref(a), it, check_iter, sigma, kkt)
        oracle_lib.oracle_check_restart(C.byref(b), it, check_iter, sigma, kkt)
        assert a.as_dict() == b.as_dict(), it
        flags.add(int(a.restart_flag))
        if a.restart_flag > 0:  # what do_restart! does to the counters
            for r in (a, b):
                r.last_gap, r.save_gap, r.times, r.inner = r.current_gap, float("inf"), r.times + 1, 0
    assert {0, 1}.issubset(flags)
