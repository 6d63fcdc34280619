"""CBF reader (hprsocp_cbf_read, host code) against the reference's CBF tests
(test/runtests.jl:167-190, the layout of data/model.cbf) and the closed form of
data/model.cbf: min -3x - 5y + z s.t. x^2 + y^2 <= z, x + 2y <= 10, 3x + y <= 12,
x, y, z >= 0, optimum (1.5, 2.5, 8.5) with objective -8.5."""
import os

import numpy as np
import pytest

import hprsocp
from kat_models import params

HERE = os.path.dirname(os.path.abspath(__file__))
MODEL = os.path.join(HERE, "golden", "model.cbf")
SQ2 = np.sqrt(2.0)


def test_model_cbf_layout(gpu_lib):
    m = hprsocp.build_from_cbf(MODEL)
    assert m.A.shape == (6, 3)
    assert m.number_eq == 0 and m.number_ineq == 2 and m.number_linear_con == 2
    np.testing.assert_array_equal(m.SOC_con_idx, [3, 7])     # 1-based, as the reference
    np.testing.assert_array_equal(m.SOC_var_idx, [4])
    np.testing.assert_allclose(m.c, [-3.0, -5.0, 1.0])
    A = m.A.toarray()
    np.testing.assert_allclose(A, [[-1, -2, 0], [-3, -1, 0], [0, 0, 0.5], [0, 0, 0.5], [SQ2, 0, 0], [0, SQ2, 0]])
    np.testing.assert_allclose(m.AL[:2], [-10.0, -12.0])     # L- rows negated: x + 2y <= 10
    assert np.all(np.isinf(m.AU))
    np.testing.assert_allclose(m.soc_rhs, [-1.0, 1.0, 0.0, 0.0])
    np.testing.assert_allclose(m.l, 0.0)
    assert np.all(np.isinf(m.u))


def test_objsense_alias(tmp_path, gpu_lib):
    text = open(MODEL).read().replace("SENSE\nmin", "OBJSENSE\nmax")
    p = tmp_path / "objsense_alias.cbf"
    p.write_text(text)
    m = hprsocp.build_from_cbf(str(p))
    np.testing.assert_allclose(m.c, [3.0, 5.0, -1.0])
    assert m.obj_constant == 0.0
    np.testing.assert_array_equal(m.SOC_con_idx, [3, 7])
    np.testing.assert_array_equal(m.SOC_var_idx, [4])


def test_cbf_errors(tmp_path, gpu_lib):
    bad = tmp_path / "x.txt"
    bad.write_text("CON\n1 1\nL= 1\n")
    with pytest.raises(hprsocp.HPRSOCPError, match="Please provide a .cbf file"):
        hprsocp.build_from_cbf(str(bad))
    p = tmp_path / "nocon.cbf"
    p.write_text("VAR\n1 1\nF 1\n")
    with pytest.raises(hprsocp.HPRSOCPError, match="missing CON"):
        hprsocp.build_from_cbf(str(p))
    p = tmp_path / "psd.cbf"
    p.write_text("VAR\n1 1\nF 1\nCON\n1 1\nL= 1\nPSDCON\n1\n2\n")
    with pytest.raises(hprsocp.HPRSOCPError, match="PSD constraints are not supported"):
        hprsocp.build_from_cbf(str(p))
    p = tmp_path / "cone.cbf"
    p.write_text("VAR\n1 1\nQR 1\nCON\n1 1\nL= 1\n")
    with pytest.raises(hprsocp.HPRSOCPError, match="Unsupported variable cone type QR"):
        hprsocp.build_from_cbf(str(p))


def test_cbf_free_rows_dropped_and_equalities_first(tmp_path, gpu_lib):
    # rows: F (dropped), L+ , L= ; variables: L- then F
    p = tmp_path / "mix.cbf"
    p.write_text("VAR\n2 2\nL- 1\nF 1\nCON\n3 3\nF 1\nL+ 1\nL= 1\n"
                 "ACOORD\n4\n0 0 9.0\n1 0 1.0\n2 1 2.0\n2 0 0.0\nBCOORD\n2\n1 -3.0\n2 4.0\n"
                 "OBJACOORD\n1\n1 1.5\nOBJBCOORD\n2.5\n")
    m = hprsocp.build_from_cbf(str(p))
    assert m.A.shape == (2, 2) and m.number_eq == 1 and m.number_ineq == 1
    np.testing.assert_allclose(m.A.toarray(), [[0.0, 2.0], [1.0, 0.0]])   # L= row first
    np.testing.assert_allclose(m.AL, [-4.0, 3.0])
    np.testing.assert_allclose(m.AU, [-4.0, np.inf])
    np.testing.assert_allclose(m.l, [-np.inf, -np.inf])
    np.testing.assert_allclose(m.u, [0.0, np.inf])
    np.testing.assert_allclose(m.c, [0.0, 1.5])
    assert m.obj_constant == 2.5


def test_model_cbf_oracle_solve(oracle_lib, gpu_lib):
    m = hprsocp.build_from_cbf(MODEL)
    with hprsocp.Solver(m, params(stoptol=1e-9), lib=oracle_lib, prefix="oracle_") as o:
        o.estimate_lambda()
        r = o.solve()
    assert r.status == "OPTIMAL"
    np.testing.assert_allclose(r.x, [1.5, 2.5, 8.5], atol=1e-6)
    assert r.primal_obj == pytest.approx(-8.5, abs=1e-6)


@pytest.mark.gpu
def test_model_cbf_gpu_solve(gpu_lib):
    m = hprsocp.build_from_cbf(MODEL)
    r = hprsocp.optimize(m, params(stoptol=1e-9))
    assert r.status == "OPTIMAL"
    np.testing.assert_allclose(r.x, [1.5, 2.5, 8.5], atol=1e-6)
    assert r.primal_obj == pytest.approx(-8.5, abs=1e-6)
