"""Pins the CPU oracle (the checker of the HIP backend) — CPU only.

* the reference's own known-answer tests (tests/golden/reference_kats.json);
* closed forms (data/model.cbf, the Clarabel and JuMP tests, the trust-region
  secular equation);
* mathematics: SpMV vs scipy, Moreau decomposition / idempotence of the SOC
  projection, golden section vs a dense scan, power iteration vs eigvalsh,
  scaling bookkeeping (scaled data == D_r^-1 A D_c^-1 etc.), Ruiz fixed point,
  one shared factor per cone block, and an independent KKT certificate.
"""
import ctypes as C
import re

import numpy as np
import pytest
import scipy.sparse as sp

import hprsocp
from hprsocp import synthetic
from kat_models import kats, model_from, params, trust_region_optimum
from kkt import certificate, soc_proj

K = kats()


def osolver(lib, model, p=None):
    return hprsocp.Solver(model, p or params(), lib=lib, prefix="oracle_")


# ------------------------------------------------------------- reference KATs
def test_kat_soc_variable_projection(oracle_lib):
    s = K["soc_var_projection"]
    with osolver(oracle_lib, model_from(s), params(sigma=1.0)) as o:
        o.set_scalars(sigma=1.0, lambda_A=0.0, lambda_Q=0.0)
        o.set("x", s["x"])
        o.update(0, True)
        np.testing.assert_allclose(o.get("x_bar"), s["expect_x_bar"], atol=1e-14)
        np.testing.assert_allclose(o.get("z_bar"), s["expect_z_bar"], atol=1e-14)


def test_kat_soc_constraint_projection(oracle_lib):
    s = K["soc_con_projection"]
    with osolver(oracle_lib, model_from(s), params(sigma=1.0)) as o:
        o.set_scalars(sigma=1.0, lambda_A=1.0, lambda_Q=0.0)
        o.set("x", s["x_hat"])  # unconstrained box + A'y = 0 => x_hat = x
        o.update(0, True)
        np.testing.assert_allclose(o.get("x_hat"), s["x_hat"], atol=1e-14)
        np.testing.assert_allclose(o.get("s"), s["expect_s"], atol=1e-14)
        np.testing.assert_allclose(o.get("y_bar"), s["expect_y_bar"], atol=1e-14)


def test_kat_bounds_violation(oracle_lib):
    s = K["bounds_violation"]
    spec = dict(n=3, soc_var_idx0=s["soc_var_idx0"])
    with osolver(oracle_lib, model_from(spec), params(sigma=1.0)) as o:
        o.set_scalars(sigma=1.0, lambda_A=0.0, lambda_Q=0.0)
        o.set("x_bar", s["x_bar"])
        r = o.residuals(0)
        assert r["err_Rp"] == pytest.approx(s["expect_err_Rp"], abs=1e-14)


def test_kat_soc_rp(oracle_lib):
    s = K["soc_rp"]
    with osolver(oracle_lib, model_from(s), params(sigma=1.0)) as o:
        o.set_scalars(sigma=1.0, lambda_A=1.0, lambda_Q=0.0)
        o.set("x_bar", s["x_bar"])
        o.residuals(1)
        np.testing.assert_allclose(o.get("Rp"), s["expect_Rp"], atol=1e-14)


@pytest.mark.parametrize("name", ["split_mixed", "split_linear_only", "split_soc_only"])
def test_kat_rp_split(oracle_lib, name):
    s = K[name]
    with osolver(oracle_lib, model_from(s), params(sigma=1.0)) as o:
        o.set_scalars(sigma=1.0, lambda_A=1.0, lambda_Q=0.0)
        o.set("x_bar", s["x_bar"])
        r = o.residuals(1)
        rp, ax = o.get("Rp"), o.get("Ax")
        m1 = o.model.number_linear_con
        denom = 1.0 + max(o.scalars()["norm_b_org"], np.max(np.abs(ax)))
        lin = np.max(np.abs(rp[:m1]), initial=0.0) / denom
        soc = np.max(np.abs(rp[m1:]), initial=0.0) / denom
        assert r["err_Rp_linear"] == pytest.approx(lin, abs=1e-15)
        assert r["err_Rp_soc"] == pytest.approx(soc, abs=1e-15)
        assert r["err_Rp"] == pytest.approx(max(lin, soc), abs=1e-15)
        if name == "split_mixed":
            assert r["err_Rp_linear"] > r["err_Rp_soc"] > 0.0
        elif name == "split_linear_only":
            assert r["err_Rp_soc"] == 0.0 and r["err_Rp_linear"] == r["err_Rp"]
        else:
            assert r["err_Rp_linear"] == 0.0 and r["err_Rp_soc"] == r["err_Rp"]


def test_kat_dual_objective_rhs_shift(oracle_lib):
    s = K["dual_objective_rhs_shift"]
    with osolver(oracle_lib, model_from(s)) as o:
        o.set_scalars(lambda_A=1.0, lambda_Q=0.0)
        o.set("y_bar", s["y_bar"])
        r = o.residuals(0)
        assert r["dual_obj"] == pytest.approx(s["expect_dual_obj"], abs=1e-10)


def test_kat_mnorm_nonnegative(oracle_lib):
    s = K["mnorm_nonnegative"]
    with osolver(oracle_lib, model_from(s), params(sigma=1.0)) as o:
        o.set_scalars(sigma=1.0, lambda_A=0.0, lambda_Q=1.0)
        o.set("dw", s["dw"])
        v = o.mnorm()
        assert np.isfinite(v) and v >= 0.0


# ------------------------------------------------------------- end-to-end KATs
def _solve(lib, model, **kw):
    with osolver(lib, model, params(**kw)) as o:
        o.estimate_lambda()
        return o.solve()


@pytest.mark.parametrize("name", ["solve_soc_variables_only", "solve_pure_soc_constraints",
                                  "solve_mixed_linear_soc"])
def test_kat_small_solves(oracle_lib, name):
    s = K[name]
    r = _solve(oracle_lib, model_from(s), stoptol=s["stoptol"], max_iter=s["max_iter"])
    assert r.status == "OPTIMAL"
    assert np.linalg.norm(r.x) <= 1e-3
    if name == "solve_soc_variables_only":
        assert r.x[0] + 1e-8 >= np.linalg.norm(r.x[1:3])
    if name == "solve_mixed_linear_soc":
        assert r.x[0] >= -1e-8


def test_kat_trust_region(oracle_lib):
    s = K["trust_region"]
    model = model_from(s)
    r = _solve(oracle_lib, model, stoptol=s["stoptol"], max_iter=s["max_iter"])
    assert r.status == "OPTIMAL"
    A = np.array(s["A"])
    assert np.linalg.norm(A @ r.x - np.array(s["rhs"])) <= 1e-5
    assert r.x[0] + 1e-8 >= np.linalg.norm(r.x[1:])
    xs = trust_region_optimum(s["Q_diag"], s["c"])
    np.testing.assert_allclose(r.x, xs, atol=2e-3)


@pytest.mark.parametrize("name", ["clarabel_var_cone", "clarabel_con_cone", "jump_soc"])
def test_kat_closed_forms(oracle_lib, name):
    s = K[name]
    r = _solve(oracle_lib, model_from(s), stoptol=s["stoptol"], max_iter=s.get("max_iter", 4000))
    assert r.status == "OPTIMAL"
    np.testing.assert_allclose(r.x, s["expect_x"], atol=s["atol"])


def test_model_cbf_closed_form(oracle_lib):
    s = K["model_cbf"]
    model = model_from(s)
    r = _solve(oracle_lib, model, stoptol=1e-8)
    assert r.status == "OPTIMAL"
    np.testing.assert_allclose(r.x, s["expect_x"], atol=1e-5)
    assert r.primal_obj == pytest.approx(s["expect_obj"], abs=1e-6)
    cert = certificate(model, r.x, r.y, r.z)
    assert cert["err_Rp"] < 1e-7 and cert["err_Rd"] < 1e-7 and cert["rel_gap"] < 1e-6


def test_builder_metadata():
    s = K["builder_metadata"]
    m = hprsocp.build_from_SOCP_data(sp.csr_matrix((4, 4)), np.zeros(4), sp.eye(4), s["rhs"],
                                     s["SOC_con_idx"], s["number_eq"], s["number_ineq"],
                                     np.full(4, -np.inf), np.full(4, np.inf), s["SOC_var_idx"])
    e = s["expect"]
    assert m.number_linear_con == e["number_linear_con"]
    assert m.number_lu_x == e["number_lu_x"]
    np.testing.assert_array_equal(m.soc_rhs, e["soc_rhs"])
    assert m.AL[0] == e["AL0"] and m.AU[0] == e["AU0"] and m.AL[1] == e["AL1"]
    assert np.isinf(m.AU[1]) and np.isinf(m.AL[2]) and np.isinf(m.AU[2])


def test_builder_errors():
    Q = sp.csr_matrix((2, 2))
    A = sp.eye(2)
    ok = dict(Q=Q, c=np.zeros(2), A=A, rhs=np.zeros(2), SOC_con_idx=[3], number_eq=1,
              number_ineq=1, l=np.zeros(2), u=np.ones(2), SOC_var_idx=[3])
    hprsocp.build_from_SOCP_data(**ok)
    for key, val, msg in [("c", np.zeros(3), "length(c)"), ("rhs", np.zeros(1), "length(rhs)"),
                          ("SOC_con_idx", [2], "must start"), ("SOC_var_idx", [2], "must end"),
                          ("number_eq", -1, "nonnegative")]:
        bad = dict(ok)
        bad[key] = val
        with pytest.raises(ValueError, match=re.escape(msg)):
            hprsocp.build_from_SOCP_data(**bad)


# ------------------------------------------------------------- mathematics
def test_soc_projection_moreau_and_idempotence(oracle_lib):
    rng = np.random.default_rng(3)
    fn = oracle_lib.oracle_soc_project
    fn.restype = None
    fn.argtypes = [C.POINTER(C.c_double), C.c_int32]
    for trial in range(200):
        k = int(rng.integers(1, 12))
        v = rng.standard_normal(k) * rng.uniform(0.1, 10)
        if trial % 3 == 0:
            v[0] = np.linalg.norm(v[1:]) * rng.choice([-1.0, 1.0])  # boundary cases
        p = v.copy()
        fn(p.ctypes.data_as(C.POINTER(C.c_double)), k)
        q = -v.copy()
        fn(q.ctypes.data_as(C.POINTER(C.c_double)), k)
        np.testing.assert_allclose(p - q, v, atol=1e-12 * (1 + np.abs(v).max()))
        assert abs(p @ q) <= 1e-10 * (1 + v @ v)
        pp = p.copy()
        fn(pp.ctypes.data_as(C.POINTER(C.c_double)), k)
        np.testing.assert_allclose(pp, p, atol=1e-12 * (1 + np.abs(p).max()))
        np.testing.assert_allclose(p, soc_proj(v), atol=1e-12 * (1 + np.abs(v).max()))


def test_golden_section_vs_scan(oracle_lib):
    g = oracle_lib.oracle_golden
    g.restype = C.c_double
    g.argtypes = [C.c_double] * 7
    rng = np.random.default_rng(5)
    for _ in range(20):
        a, b, c, d = rng.uniform(0.01, 10, 4)
        x = g(a, b, c, d, 1e-12, 1e12, 1e-13)
        f = lambda t: a * t + b / t + c * t * t / (1 + d * t)
        grid = np.logspace(-6, 6, 200001)
        xs = grid[np.argmin(f(grid))]
        assert f(x) <= f(xs) * (1 + 1e-9)
        assert abs(np.log(x / xs)) < 1e-3


def test_spmv_vs_scipy(oracle_lib):
    model = synthetic.make_c0(seed=11)
    with osolver(oracle_lib, model) as o:
        Av = o.scaled_A()
        A = model.A.tocsr()
        As = sp.csr_matrix((Av, A.indices, A.indptr), shape=A.shape)
        v = np.random.default_rng(1).standard_normal(A.shape[1])
        w = np.random.default_rng(2).standard_normal(A.shape[0])
        np.testing.assert_allclose(o.spmv(v), As @ v, rtol=1e-13, atol=1e-13)
        np.testing.assert_allclose(o.spmv(w, trans=True), As.T @ w, rtol=1e-13, atol=1e-13)


def test_power_iteration_vs_eigvalsh(oracle_lib):
    model = synthetic.make_c0(seed=12)
    with osolver(oracle_lib, model) as o:
        sc = o.estimate_lambda()
        A = model.A.tocsr()
        As = sp.csr_matrix((o.scaled_A(), A.indices, A.indptr), shape=A.shape).toarray()
        lam = np.linalg.eigvalsh(As @ As.T).max()
        assert sc["lambda_max_A"] / 1.05 == pytest.approx(lam, rel=1e-3)


def test_scaling_bookkeeping(oracle_lib):
    """Scaled data equals D_r^-1 A D_c^-1, c/(D_c c_s), bounds/(D_r b_s), l D_c/b_s;
    every cone block shares one row/column factor (scaling.jl:500-526)."""
    model = synthetic.make_small_qp(seed=3, q="none")
    with osolver(oracle_lib, model) as o:
        rn, cn = o.get("row_norm"), o.get("col_norm")
        sc = o.scalars()
        A = model.A.tocsr()
        rows = np.repeat(np.arange(A.shape[0]), np.diff(A.indptr))
        expect = A.data / rn[rows] / cn[A.indices]
        np.testing.assert_allclose(o.scaled_A(), expect, rtol=1e-12)
        np.testing.assert_allclose(o.get("c"), model.c / cn / sc["c_scale"], rtol=1e-12)
        fin = np.isfinite(model.AL)
        np.testing.assert_allclose(o.get("AL")[fin], model.AL[fin] / rn[fin] / sc["b_scale"], rtol=1e-12)
        fin = np.isfinite(model.l)
        np.testing.assert_allclose(o.get("l")[fin], model.l[fin] * cn[fin] / sc["b_scale"], rtol=1e-12)
        ci = np.asarray(model.SOC_con_idx) - 1
        for k in range(len(ci) - 1):
            assert np.ptp(rn[ci[k]:ci[k + 1]]) == 0.0
        vi = np.asarray(model.SOC_var_idx) - 1
        for k in range(len(vi) - 1):
            assert np.ptp(cn[vi[k]:vi[k + 1]]) == 0.0
        assert np.all(rn >= 1e-4 ** 11) and np.all(cn > 0)


def test_ruiz_fixed_point(oracle_lib):
    """Ruiz alone (max strategy, 30 passes) drives every row and column inf-norm to 1."""
    model = synthetic.make_small_qp(seed=4, q="none", var_sizes=(), con_sizes=())
    p = params(use_Pock_Chambolle_scaling=False, use_bc_scaling=False, ruiz_iterations=30,
               soc_block_scaling_strategy="max")
    with osolver(oracle_lib, model, p) as o:
        A = model.A.tocsr()
        As = sp.csr_matrix((o.scaled_A(), A.indices, A.indptr), shape=A.shape)
        rmax = np.abs(As).max(axis=1).toarray().ravel()
        cmax = np.abs(As).max(axis=0).toarray().ravel()
        nz_r = np.diff(A.indptr) > 0
        nz_c = np.bincount(A.indices, minlength=A.shape[1]) > 0
        np.testing.assert_allclose(rmax[nz_r], 1.0, rtol=1e-3)
        np.testing.assert_allclose(cmax[nz_c], 1.0, rtol=1e-3)


@pytest.mark.parametrize("q", ["none", "diag", "csr", "factored"])
def test_oracle_small_solve_certificate(oracle_lib, q):
    model = synthetic.make_small_qp(seed=21, q=q)
    r = _solve(oracle_lib, model, stoptol=1e-7, max_iter=200000)
    assert r.status == "OPTIMAL"
    cert = certificate(model, r.x, r.y, r.z)
    assert cert["err_Rp"] < 1e-6 and cert["err_Rd"] < 1e-6
    assert cert["rel_gap"] < 1e-5
    assert cert["y_cone_violation"] < 1e-5 and cert["z_cone_violation"] < 1e-5


def test_oracle_c0_certificate(oracle_lib):
    model = synthetic.make_c0(seed=0)
    r = _solve(oracle_lib, model, stoptol=1e-6, max_iter=200000)
    assert r.status == "OPTIMAL"
    cert = certificate(model, r.x, r.y, r.z)
    assert cert["err_Rp"] < 1e-5 and cert["err_Rd"] < 1e-5 and cert["rel_gap"] < 1e-5


@pytest.mark.parametrize("strategy", sorted(hprsocp._abi.SOC_STRATEGIES))
def test_oracle_soc_block_rules(oracle_lib, strategy):
    """Every SOC block rule after one Ruiz pass (no bc, no Pock-Chambolle): each
    constraint cone gets one row factor.  Without the two extra max passes of
    the pre-Pock strategies (scaling.jl:136-146, :1008-1040) that factor lies
    between the smallest and the largest factor its rows would get alone (sqrt
    of the row inf-norm) — the rules (scaling.jl:253-477) all aggregate
    max/rms-type statistics of the block — and equals the largest for the max
    rule.  Then the oracle solves the model with the full pipeline to a
    certified 1e-6."""
    model = synthetic.make_small_qp(seed=5, q="diag", var_sizes=(3, 12, 40), m1=10, con_sizes=(4, 30, 9),
                                    density=0.5)
    A = model.A.tocsr()
    rmax = np.abs(A).max(axis=1).toarray().ravel()
    assert np.all(rmax > 0)
    own = np.sqrt(rmax)
    p = params(soc_block_scaling_strategy=strategy, ruiz_iterations=1, use_bc_scaling=False,
               use_Pock_Chambolle_scaling=False)
    with osolver(oracle_lib, model, p) as o:
        rn = o.get("row_norm")
    ci = np.asarray(model.SOC_con_idx) - 1
    for k in range(len(ci) - 1):
        blk = slice(ci[k], ci[k + 1])
        assert np.ptp(rn[blk]) == 0.0
        f = rn[ci[k]]
        if 10 <= hprsocp._abi.SOC_STRATEGIES[strategy] <= 18:  # pre-Pock stage: extra passes
            continue
        assert own[blk].min() * (1 - 1e-12) <= f <= own[blk].max() * (1 + 1e-12), (k, f)
        if strategy == "max":
            assert f == pytest.approx(own[blk].max(), rel=1e-14)
    r = _solve(oracle_lib, model, soc_block_scaling_strategy=strategy, stoptol=1e-6, max_iter=200000)
    assert r.status == "OPTIMAL"
    cert = certificate(model, r.x, r.y, r.z)
    assert cert["err_Rp"] < 1e-5 and cert["err_Rd"] < 1e-5
