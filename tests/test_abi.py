"""The C-ABI library loads and exports every entry point include/hprsocp.h
declares (no compute calls: this runs without a GPU), and the ctypes mirror
matches the header's struct layouts."""
import ctypes as C
import os
import re

import pytest

import hprsocp
from hprsocp import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hprsocp.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hprsocp_[A-Za-z_0-9]+)\s*\(", text)))


def test_header_declares_abi_functions():
    names = declared_functions()
    for fn in _abi.ABI_FUNCTIONS:
        assert "hprsocp_" + fn in names
    assert "hprsocp_time_iterations" in names
    for fn in _abi._LOOP_SIGS:
        assert "hprsocp_" + fn in names


def test_restart_info_layout_matches_julia():
    """hprsocp_restart_info has the C layout of the reference's HPRSOCP_restart
    (structs.jl:600-623; Int = 8 bytes, Bool = 1 byte, C alignment), so a Julia
    host passes its own restart_info by reference."""
    julia_fields = ["restart_flag", "first_restart", "last_gap", "current_gap", "save_gap", "inner",
                    "step", "sufficient", "necessary", "long", "ratio", "times", "weighted_norm",
                    "best_gap", "best_kkt", "best_sigma", "best_iter", "sigma_correction_active",
                    "sigma_correction_hold", "sigma_correction_dir"]
    bools = {"first_restart", "sigma_correction_active"}
    off = 0
    for (name, ctype), jname in zip(_abi.RestartInfo._fields_, julia_fields):
        assert name == (jname if jname != "long" else "long_restarts")
        size = 1 if jname in bools else 8
        off = (off + size - 1) // size * size
        assert getattr(_abi.RestartInfo, name).offset == off, name
        assert C.sizeof(ctype) == size, name
        off += size
    assert C.sizeof(_abi.RestartInfo) == (off + 7) // 8 * 8


def test_library_exports_every_declared_symbol():
    path = _abi.library_path("libhprsocp.so")
    if not os.path.exists(path):
        pytest.skip("libhprsocp.so not built (run __graft_entry__.build())")
    lib = C.CDLL(path)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"missing exports: {missing}"
    hprsocp.load_library()  # binds every signature


def test_oracle_exports_mirror(oracle_lib):
    for fn in _abi.ABI_FUNCTIONS:
        assert hasattr(oracle_lib, "oracle_" + fn)


def test_default_params_match_reference(oracle_lib):
    # HPRSOCP_parameters() defaults, reference src/structs.jl:277
    p = _abi.Params()
    oracle_lib.oracle_default_params(C.byref(p))
    py = hprsocp.HPRSOCP_parameters()
    cp = py.to_c([])
    for name, _ in _abi.Params._fields_:
        if name in ("initial_x", "initial_y", "reserved"):
            continue
        assert getattr(p, name) == getattr(cp, name), name
    assert (p.stoptol, p.sigma, p.max_iter, p.check_iter, p.ruiz_iterations) == (1e-6, -1.0, 2**31 - 1, 150, 10)
    assert p.soc_block_scaling_strategy == _abi.SOC_STRATEGIES["phase_taper"]


def test_struct_sizes(oracle_lib):
    # the ctypes mirror has the C compiler's layout of every ABI struct
    fn = oracle_lib.oracle_abi_sizeof
    fn.restype = C.c_int64
    fn.argtypes = [C.c_int32]
    for k, st in enumerate([_abi.Problem, _abi.Params, _abi.Result, _abi.Residuals, _abi.Scalars,
                            _abi.RestartInfo]):
        assert C.sizeof(st) == fn(k), st.__name__


def test_invalid_problem_rejected_by_oracle(oracle_lib):
    """Error behaviour of create (builders.jl / allocate_workspace checks)."""
    import numpy as np
    import scipy.sparse as sp
    model = hprsocp.build_qp(sp.eye(2), np.zeros(2), [0, 0], [1, 1], [-1, -1], [1, 1], [], 2, 2,
                             [2], [2])
    with pytest.raises(hprsocp.HPRSOCPError):
        hprsocp.Solver(model, hprsocp.HPRSOCP_parameters(sigma=-3.0, verbose=False),
                       lib=oracle_lib, prefix="oracle_")
    bad = hprsocp.build_qp(sp.eye(2), np.zeros(2), [0, 0], [1, 1], [-1, -1], [1, 1], [], 2, 2,
                           [1], [2])
    with pytest.raises(hprsocp.HPRSOCPError):
        hprsocp.Solver(bad, hprsocp.HPRSOCP_parameters(verbose=False), lib=oracle_lib, prefix="oracle_")
