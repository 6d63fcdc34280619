"""The Julia binding julia/HPRSOCPHip.jl (module HPRSOCPHip) passes
hprsocp_problem, hprsocp_params, hprsocp_result and hprsocp_residuals to ccall
as isbits structs that repeat the C layout field by field (Julia lays out
isbits structs with C alignment rules).  No Julia toolchain exists here, so the
shim itself cannot run; this test parses its struct definitions and checks
them against the C layout: the same fields in the same order at the same
offsets, and the same size, as the ctypes mirror (whose sizes test_abi checks
against the compiled library).  It also checks that a Q operator without a
factored form is rejected, never passed on as an empty Q."""
import ctypes as C
import os
import re
import sys

import pytest

from hprsocp import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DOC = os.path.join(ROOT, "julia", "HPRSOCPHip.jl")

# Julia type -> (size, alignment) on x86-64 / aarch64 Linux
JULIA_TYPES = {"Cint": (4, 4), "Int32": (4, 4), "Int64": (8, 8), "Cdouble": (8, 8), "Float64": (8, 8),
               "Cstring": (8, 8)}


def julia_type(t):
    t = t.strip()
    if t.startswith("Ptr{"):
        return 8, 8
    m = re.fullmatch(r"NTuple\{(\d+),\s*(\w+)\}", t)
    if m:
        size, align = JULIA_TYPES[m.group(2)]
        return int(m.group(1)) * size, align
    return JULIA_TYPES[t]


def julia_struct(name):
    """[(field, offset)], size of `struct name` / `mutable struct name` in the doc."""
    text = open(DOC).read()
    m = re.search(r"^(?:mutable )?struct " + name + r"\b.*?\n(.*?)^end", text, flags=re.S | re.M)
    assert m, name
    fields, off, align_max = [], 0, 1
    for line in m.group(1).splitlines():
        line = line.split("#")[0]
        for decl in line.split(";"):
            decl = decl.strip()
            if "::" not in decl or "(" in decl.split("::")[0]:
                continue
            fname, ftype = decl.split("::", 1)
            size, align = julia_type(ftype)
            off = (off + align - 1) // align * align
            fields.append((fname.strip(), off))
            off += size
            align_max = max(align_max, align)
    return fields, (off + align_max - 1) // align_max * align_max


def c_struct(cls):
    return [(f[0], getattr(cls, f[0]).offset) for f in cls._fields_], C.sizeof(cls)


@pytest.mark.parametrize("jname,cls", [("CProblem", _abi.Problem), ("CParams", _abi.Params),
                                       ("CResult", _abi.Result), ("CResiduals", _abi.Residuals)])
def test_julia_structs_match_the_c_layout(jname, cls):
    jf, jsize = julia_struct(jname)
    cf, csize = c_struct(cls)
    assert [f for f, _ in jf] == [f for f, _ in cf]
    assert jf == cf
    assert jsize == csize


def test_parser_sees_a_changed_field(tmp_path, monkeypatch):
    """The check fails when the notes' struct drifts: a copy without
    CParams.save_filename no longer matches the C layout."""
    text = open(DOC).read()
    assert "auto_save::Cint; save_filename::Cstring;" in text
    bad = tmp_path / "HPRSOCPHip.jl"
    bad.write_text(text.replace("auto_save::Cint; save_filename::Cstring;", "auto_save::Cint;"))
    monkeypatch.setattr(sys.modules[__name__], "DOC", str(bad))
    jf, jsize = julia_struct("CParams")
    cf, csize = c_struct(_abi.Params)
    assert jf != cf and jsize != csize


def test_operator_q_is_factored_or_rejected():
    """An AbstractQOperatorCPU Q goes to the backend only in the factored form
    F'F + Diagonal(d) (HPRSOCP_Q_FACTORED); without a hip_factored method the
    shim raises an error.  A SparseMatrixCSC Q goes as CSR; nothing else is
    accepted."""
    text = open(DOC).read()
    body = text[text.index("function problem_arrays"):text.index("ptr_or_null(v::Vector)")]
    sparse_branch, rest = body.split("elseif Q isa AbstractQOperatorCPU", 1)
    op_branch, other = rest.split("else", 1)
    assert "if Q isa SparseMatrixCSC" in sparse_branch and "HPRSOCP_Q_CSR" in sparse_branch
    assert "hasmethod(hip_factored" in op_branch and "error(" in op_branch
    assert "HPRSOCP_Q_FACTORED" in op_branch and "hip_factored(Q)" in op_branch
    assert other.lstrip().startswith('error("HPRSOCPHip: unsupported Q type')
    assert "const HPRSOCP_Q_CSR = Cint(0)" in text and "const HPRSOCP_Q_FACTORED = Cint(1)" in text


def test_batch_binding_uses_contiguous_results():
    """optimize_batch_hip passes hprsocp_batch_solve a Vector{CResult}, which is one
    contiguous C array only when CResult is an isbits (immutable) struct; it
    creates, solves, reads every instance and destroys the batch."""
    text = open(DOC).read()
    assert re.search(r"^struct CResult\b", text, flags=re.M), "CResult must be immutable (isbits)"
    body = text[text.index("function optimize_batch_hip"):text.index("# ---------------------------------------------------------------- seams")]
    for sym in ("hprsocp_batch_create", "hprsocp_batch_estimate_lambda", "hprsocp_batch_set_initial",
                "hprsocp_batch_solve", "hprsocp_batch_get_solution", "hprsocp_batch_destroy"):
        assert f":{sym}" in body, sym
    assert "Vector{CResult}(undef, k)" in body and "Ptr{CResult}" in body
    assert "finally" in body  # the batch is destroyed on every path
