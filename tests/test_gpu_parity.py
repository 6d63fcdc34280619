"""HIP backend vs the CPU oracle, through the C ABI, on the same seeded inputs.

Tolerances (FP64 everywhere; only summation order differs):
  * scaled data, scale factors, sigma_0:         rtol 1e-12
  * lambda_max(AA^T), lambda_max(Q):             rtol 1e-8
  * iterates after 30 HPR steps from one state:   |a-b|_inf <= 1e-9 (1 + |b|_inf)
  * residual block, M-norm:                       rtol 1e-9
  * end-to-end: GPU reaches KKT < stoptol = 1e-6, its independent certificate is
    < 1e-5, and its primal objective agrees with an FP64 oracle solve at 1e-9 to
    1e-6 relative; iteration, milestone and restart counts equal the oracle's,
    final sigma to 1e-9 relative (1e-7 for QP, whose sigma comes from a
    golden-section search).
"""
import numpy as np
import pytest
import scipy.sparse as sp

import hprsocp
from hprsocp import synthetic
from kat_models import kats, model_from, params, trust_region_optimum
from kkt import certificate

pytestmark = pytest.mark.gpu

MODELS = {
    "c0": lambda: synthetic.make_c0(seed=0),
    "noq_small": lambda: synthetic.make_small_qp(seed=3, q="none"),
    "qp_diag": lambda: synthetic.make_small_qp(seed=7, q="diag"),
    "qp_csr": lambda: synthetic.make_small_qp(seed=8, q="csr"),
    "qp_factored": lambda: synthetic.make_small_qp(seed=9, q="factored"),
    # cones longer than one 256-row tile on both sides (two-phase projection) and
    # rows long enough for the whole-workgroup SpMV path
    "huge_cones": lambda: synthetic.make_small_qp(seed=10, n1=300, var_sizes=(700, 3, 1, 300),
                                                  m1=40, con_sizes=(900, 5, 2), density=0.02,
                                                  q="diag"),
    "huge_noq": lambda: synthetic.make_small_qp(seed=11, n1=50, var_sizes=(1500,), m1=5,
                                                con_sizes=(600,), density=0.6, q="none"),
    # a constraint cone of 5000 rows: the residual's cone statistics sum it in
    # three pieces of at most 2048 rows
    "long_cone": lambda: synthetic.make_small_qp(seed=12, n1=120, var_sizes=(3, 5), m1=20,
                                                 con_sizes=(5000, 4), density=0.01, q="none"),
    "c1_tiny": lambda: synthetic.make_c1(scale=0.002),
    "c2_tiny": lambda: synthetic.make_c2(scale=0.004),
}

_cache = {}


def model(name):
    if name not in _cache:
        _cache[name] = MODELS[name]()
    return _cache[name]


def pair(name, oracle_lib, gpu_lib, **kw):
    m = model(name)
    p = params(**kw)
    print(f"[case] {name} m={m.A.shape[0]} n={m.A.shape[1]} nnz={m.A.nnz}", flush=True)
    g = hprsocp.Solver(m, p)
    o = hprsocp.Solver(m, p, lib=oracle_lib, prefix="oracle_")
    return m, g, o


def close(a, b, tol):
    a, b = np.asarray(a), np.asarray(b)
    fa, fb = np.isfinite(a), np.isfinite(b)
    assert np.array_equal(fa, fb)
    assert np.array_equal(a[~fa], b[~fb])
    if fa.any():
        err = np.max(np.abs(a[fa] - b[fb]))
        assert err <= tol * (1 + np.max(np.abs(b[fb]))), f"max abs err {err}"


@pytest.mark.parametrize("name", list(MODELS))
def test_scaling_parity(name, oracle_lib, gpu_lib):
    m, g, o = pair(name, oracle_lib, gpu_lib)
    with g, o:
        sg, so = g.scalars(), o.scalars()
        for k in ("b_scale", "c_scale", "norm_b", "norm_c", "norm_b_org", "norm_c_org", "sigma"):
            assert sg[k] == pytest.approx(so[k], rel=1e-12, abs=1e-300), k
        assert sg["Q_is_diag"] == so["Q_is_diag"] and sg["has_quadratic"] == so["has_quadratic"]
        for v in ("c", "l", "u", "AL", "AU", "soc_rhs_full", "row_norm", "col_norm", "diag_Q"):
            close(g.get(v), o.get(v), 1e-12)
        close(g.scaled_A(), o.scaled_A(), 1e-12)


SCALING_OPTIONS = [dict(soc_block_scaling_strategy=k) for k in sorted(hprsocp._abi.SOC_STRATEGIES)] + [
    dict(bc_scaling_norm_type="rms"), dict(bc_scaling_norm_type="linf"),
    dict(use_Ruiz_scaling=False), dict(use_Pock_Chambolle_scaling=False), dict(use_bc_scaling=False),
    dict(ruiz_iterations=3, use_bc_scaling=False, use_Pock_Chambolle_scaling=False),
]


@pytest.mark.parametrize("opts", SCALING_OPTIONS, ids=lambda o: "-".join(f"{k}={v}" for k, v in o.items()))
@pytest.mark.parametrize("name", ["huge_cones", "c2_tiny"])
def test_scaling_options_parity(name, opts, oracle_lib, gpu_lib):
    """Every SOC block rule (scaling.jl:253-477), every bc norm and every switch
    of the scaling pipeline: the GPU's scaled problem equals the oracle's
    (two independent restatements) to 1e-12, on a model with long and short
    cones of both kinds and on one with 20k size-3 cones."""
    m, g, o = pair(name, oracle_lib, gpu_lib, **opts)
    with g, o:
        sg, so = g.scalars(), o.scalars()
        for k in ("b_scale", "c_scale", "norm_b", "norm_c", "sigma"):
            assert sg[k] == pytest.approx(so[k], rel=1e-12, abs=1e-300), k
        for v in ("row_norm", "col_norm", "c", "AL", "AU", "l", "u", "soc_rhs_full"):
            close(g.get(v), o.get(v), 1e-12)
        close(g.scaled_A(), o.scaled_A(), 1e-12)


@pytest.mark.parametrize("name", list(MODELS))
def test_lambda_parity(name, oracle_lib, gpu_lib):
    m, g, o = pair(name, oracle_lib, gpu_lib)
    with g, o:
        sg, so = g.estimate_lambda(), o.estimate_lambda()
        assert sg["lambda_max_A"] == pytest.approx(so["lambda_max_A"], rel=1e-8)
        assert sg["lambda_max_Q"] == pytest.approx(so["lambda_max_Q"], rel=1e-8)


@pytest.mark.parametrize("name", list(MODELS))
def test_spmv_vs_scipy(name, oracle_lib, gpu_lib):
    m, g, o = pair(name, oracle_lib, gpu_lib)
    with g, o:
        A = m.A.tocsr()
        As = sp.csr_matrix((g.scaled_A(), A.indices, A.indptr), shape=A.shape)
        rng = np.random.default_rng(0)
        v, w = rng.standard_normal(A.shape[1]), rng.standard_normal(A.shape[0])
        close(g.spmv(v), As @ v, 1e-13)
        close(g.spmv(w, trans=True), As.T @ w, 1e-13)


STATE = {"noq": ["x", "y"], "qp": ["x", "y", "w", "ATy"]}
FULL = ["x_bar", "z_bar", "x_hat", "dx", "y_bar", "s", "dy"]


@pytest.mark.parametrize("name", list(MODELS))
def test_iteration_parity(name, oracle_lib, gpu_lib):
    """30 HPR steps (mixed full/partial) from the same state and scalars."""
    m, g, o = pair(name, oracle_lib, gpu_lib)
    with g, o:
        so = o.estimate_lambda()
        for s in (g, o):
            s.set_scalars(sigma=so["sigma"], lambda_A=so["lambda_max_A"], lambda_Q=so["lambda_max_Q"])
        kind = "qp" if so["has_quadratic"] else "noq"
        for k in range(30):
            full = (k % 7 == 6) or k == 29
            g.update(k, full)
            o.update(k, full)
        for v in STATE[kind] + FULL + (["w_bar", "dw"] if kind == "qp" else []):
            close(g.get(v), o.get(v), 1e-9)
        # residual block and M-norm on the same state
        for it in (0, 1):
            rg, ro = g.residuals(it), o.residuals(it)
            for k in ro:
                assert rg[k] == pytest.approx(ro[k], rel=1e-9, abs=1e-13), (it, k)
        assert g.mnorm() == pytest.approx(o.mnorm(), rel=1e-9, abs=1e-14)


@pytest.mark.parametrize("name", ["c0", "noq_small", "qp_diag", "qp_csr", "qp_factored",
                                  "huge_cones", "long_cone", "c1_tiny", "c2_tiny"])
def test_solve_parity(name, oracle_lib, gpu_lib):
    """Same algorithm, same stoptol: the GPU and the oracle solve to the same point
    (objective to 1e-6 relative); both are within 1e-5 of a 1e-9 oracle solve."""
    m = model(name)
    print(f"[case] solve {name}", flush=True)
    pg = params(stoptol=1e-6, max_iter=100000, time_limit=100.0)
    with hprsocp.Solver(m, pg) as g:
        g.estimate_lambda()
        rg = g.solve()
    assert rg.status == "OPTIMAL", (rg.status, rg.residuals, rg.iter)
    assert rg.residuals < 1e-6
    cert = certificate(m, rg.x, rg.y, rg.z)
    assert cert["err_Rp"] < 1e-5 and cert["err_Rd"] < 1e-5, cert
    with hprsocp.Solver(m, pg, lib=oracle_lib, prefix="oracle_") as o:
        o.estimate_lambda()
        ro = o.solve()
    assert ro.status == "OPTIMAL"
    po = params(stoptol=1e-9, max_iter=400000, time_limit=300.0)
    with hprsocp.Solver(m, po, lib=oracle_lib, prefix="oracle_") as o:
        o.estimate_lambda()
        rs = o.solve()
    rel = abs(rg.primal_obj - ro.primal_obj) / (1 + abs(ro.primal_obj))
    rel9 = abs(rg.primal_obj - rs.primal_obj) / (1 + abs(rs.primal_obj))
    print(f"[result] {name}: gpu iter {rg.iter} obj {rg.primal_obj:.12e} | oracle(1e-6) iter {ro.iter} "
          f"obj {ro.primal_obj:.12e} rel {rel:.2e} | oracle(1e-9) obj {rs.primal_obj:.12e} "
          f"rel {rel9:.2e}", flush=True)
    assert rel <= 1e-6
    assert rel9 <= 1e-5
    np.testing.assert_allclose(rg.x, o_x := ro.x, atol=1e-6 * (1 + np.abs(o_x).max()) * 1e3)
    # same FP64 algorithm on both sides (only summation order differs): the same
    # trajectory of restart decisions and sigma updates
    assert rg.iter == ro.iter, (rg.iter, ro.iter)
    assert rg.restarts == ro.restarts, (rg.restarts, ro.restarts)
    assert (rg.iter_4, rg.iter_6) == (ro.iter_4, ro.iter_6)
    # QP sigma proposals come from golden-section searches, whose minimiser the
    # dots' rounding fixes only to ~sqrt(eps) relative (the decisions above agree)
    qp = m.Q is not None and (isinstance(m.Q, tuple) or sp.csr_matrix(m.Q).nnz > 0)
    assert rg.sigma == pytest.approx(ro.sigma, rel=1e-7 if qp else 1e-9)
    assert rg.lambda_max_A == pytest.approx(ro.lambda_max_A, rel=1e-8)


# ------------------------------------------------------------ reference KATs on the GPU
K = kats()


def gsolver(spec, **kw):
    return hprsocp.Solver(model_from(spec), params(**kw))


def test_gpu_kat_soc_projections(gpu_lib):
    s = K["soc_var_projection"]
    with gsolver(s, sigma=1.0) as g:
        g.set_scalars(sigma=1.0, lambda_A=0.0, lambda_Q=0.0)
        g.set("x", s["x"])
        g.update(0, True)
        np.testing.assert_allclose(g.get("x_bar"), s["expect_x_bar"], atol=1e-14)
        np.testing.assert_allclose(g.get("z_bar"), s["expect_z_bar"], atol=1e-14)
    s = K["soc_con_projection"]
    with gsolver(s, sigma=1.0) as g:
        g.set_scalars(sigma=1.0, lambda_A=1.0, lambda_Q=0.0)
        g.set("x", s["x_hat"])
        g.update(0, True)
        np.testing.assert_allclose(g.get("s"), s["expect_s"], atol=1e-14)
        np.testing.assert_allclose(g.get("y_bar"), s["expect_y_bar"], atol=1e-14)


def test_gpu_kat_residuals(gpu_lib):
    s = K["bounds_violation"]
    with gsolver(dict(n=3, soc_var_idx0=s["soc_var_idx0"]), sigma=1.0) as g:
        g.set_scalars(sigma=1.0, lambda_A=0.0, lambda_Q=0.0)
        g.set("x_bar", s["x_bar"])
        assert g.residuals(0)["err_Rp"] == pytest.approx(0.5, abs=1e-14)
    s = K["soc_rp"]
    with gsolver(s, sigma=1.0) as g:
        g.set_scalars(sigma=1.0, lambda_A=1.0, lambda_Q=0.0)
        g.set("x_bar", s["x_bar"])
        g.residuals(1)
        np.testing.assert_allclose(g.get("Rp"), s["expect_Rp"], atol=1e-14)
    s = K["dual_objective_rhs_shift"]
    with gsolver(s) as g:
        g.set_scalars(lambda_A=1.0, lambda_Q=0.0)
        g.set("y_bar", s["y_bar"])
        assert g.residuals(0)["dual_obj"] == pytest.approx(1.0, abs=1e-10)
    s = K["mnorm_nonnegative"]
    with gsolver(s, sigma=1.0) as g:
        g.set_scalars(sigma=1.0, lambda_A=0.0, lambda_Q=1.0)
        g.set("dw", s["dw"])
        v = g.mnorm()
        assert np.isfinite(v) and v >= 0.0


@pytest.mark.parametrize("name", ["solve_soc_variables_only", "solve_pure_soc_constraints",
                                  "solve_mixed_linear_soc", "clarabel_var_cone",
                                  "clarabel_con_cone", "jump_soc", "trust_region", "model_cbf"])
@pytest.mark.parametrize("precision", ["fp64", "fp32_delta"])
def test_gpu_kat_solves(name, precision, gpu_lib):
    """The reference's small end-to-end solves (and the closed forms they pin),
    in FP64 and in the FP32 delta form, which must meet the same checks."""
    s = K[name]
    tol = s.get("stoptol", 1e-8)
    r = hprsocp.optimize(model_from(s), params(stoptol=tol, max_iter=s.get("max_iter", 100000),
                                               precision=precision))
    assert r.status == "OPTIMAL"
    if "expect_x" in s:
        np.testing.assert_allclose(r.x, s["expect_x"], atol=s.get("atol", 1e-5))
    if name == "trust_region":
        np.testing.assert_allclose(r.x, trust_region_optimum(s["Q_diag"], s["c"]), atol=2e-3)
        assert np.linalg.norm(np.array(s["A"]) @ r.x - np.array(s["rhs"])) <= 1e-5
    if name.startswith("solve_"):
        assert np.linalg.norm(r.x) <= 1e-3
    if name == "model_cbf":
        assert r.primal_obj == pytest.approx(-8.5, abs=1e-6)


def test_gpu_rejects_bad_input(gpu_lib):
    m = hprsocp.build_qp(sp.eye(2), np.zeros(2), [0, 0], [1, 1], [-1, -1], [1, 1], [], 2, 2, [1], [2])
    with pytest.raises(hprsocp.HPRSOCPError):
        hprsocp.Solver(m, params())
    with pytest.raises(hprsocp.HPRSOCPError):
        hprsocp.Solver(model("c0"), params(use_gpu=False))


# ------------------------------------------------------------ XCD-sliced SpMV path
@pytest.fixture
def forced_slices(request, monkeypatch):
    """(slices, column-sorted layout) forced on every operator of the next create."""
    P, srt = request.param
    monkeypatch.setenv("HPRSOCP_XCD_SLICES", str(P))
    monkeypatch.setenv("HPRSOCP_XCD_SORTED", str(int(srt)))
    return request.param


SLICED = ["c0", "qp_csr", "qp_factored", "huge_cones", "c1_tiny", "c2_tiny"]


@pytest.mark.parametrize("forced_slices", [(8, False), (3, False), (8, True), (3, True)], indirect=True)
@pytest.mark.parametrize("name", SLICED)
def test_sliced_spmv_and_iterations(name, forced_slices, oracle_lib, gpu_lib):
    """Column-sliced SpMV (forced on small models), row-major and column-sorted
    layouts: SpMV vs scipy, then 30 HPR steps, residuals and M-norm against the
    oracle."""
    m, g, o = pair(name, oracle_lib, gpu_lib)
    with g, o:
        A = m.A.tocsr()
        As = sp.csr_matrix((g.scaled_A(), A.indices, A.indptr), shape=A.shape)
        rng = np.random.default_rng(1)
        v, w = rng.standard_normal(A.shape[1]), rng.standard_normal(A.shape[0])
        close(g.spmv(v), As @ v, 1e-13)
        close(g.spmv(w, trans=True), As.T @ w, 1e-13)
        so = o.estimate_lambda()
        sg = g.estimate_lambda()
        assert sg["lambda_max_A"] == pytest.approx(so["lambda_max_A"], rel=1e-8)
        for s in (g, o):
            s.set_scalars(sigma=so["sigma"], lambda_A=so["lambda_max_A"], lambda_Q=so["lambda_max_Q"])
        kind = "qp" if so["has_quadratic"] else "noq"
        for k in range(30):
            full = (k % 7 == 6) or k == 29
            g.update(k, full)
            o.update(k, full)
        for vname in STATE[kind] + FULL:
            close(g.get(vname), o.get(vname), 1e-9)
        rg, ro = g.residuals(1), o.residuals(1)
        for k in ro:
            assert rg[k] == pytest.approx(ro[k], rel=1e-9, abs=1e-13), k
        assert g.mnorm() == pytest.approx(o.mnorm(), rel=1e-9, abs=1e-14)


_extra = {}


def _slice_model(name):
    """c1 at 2% scale: several slice tiles per slice; dense_rows: rows with more
    than one sorted chunk of entries in a slice (SpMV only, not a solvable model)."""
    if name not in _extra:
        if name == "c1_small":
            _extra[name] = synthetic.make_c1(scale=0.02)
        elif name == "dense_rows":
            rng = np.random.default_rng(4)
            n = 60000
            A = sp.random(300, n, density=0.002, format="lil", random_state=rng, data_rvs=rng.standard_normal)
            for r in (17, 150):
                cols = rng.choice(n, 55000, replace=False)
                A[r, cols] = rng.standard_normal(55000)
            A = A.tocsr()
            m = A.shape[0]
            _extra[name] = hprsocp.build_qp(A, rng.standard_normal(n), -np.ones(m), np.ones(m),
                                            -np.ones(n), np.ones(n), [], m, n, [m], [n])
        else:
            _extra[name] = model(name)
    return _extra[name]


def _sliced_run(name, P, srt, monkeypatch, precision="fp64", solve=True, max_iter=3000):
    monkeypatch.setenv("HPRSOCP_XCD_SLICES", str(P))
    monkeypatch.setenv("HPRSOCP_XCD_SORTED", str(int(srt)))
    m = _slice_model(name)
    with hprsocp.Solver(m, params(stoptol=1e-6, max_iter=max_iter, precision=precision)) as g:
        rng = np.random.default_rng(2)
        v, w = rng.standard_normal(m.A.shape[1]), rng.standard_normal(m.A.shape[0])
        ax, atw = g.spmv(v), g.spmv(w, trans=True)
        As = sp.csr_matrix((g.scaled_A(), m.A.indices, m.A.indptr), shape=m.A.shape)
        close(ax, As @ v, 1e-13)
        close(atw, As.T @ w, 1e-13)
        r = None
        if solve:
            g.estimate_lambda()
            r = g.solve()
    return ax, atw, r


@pytest.mark.parametrize("precision", ["fp64", "fp32_refine", "fp32_delta"])
@pytest.mark.parametrize("name", ["c1_small", "c2_tiny", "qp_factored"])
def test_sorted_slices_bit_identical(name, precision, monkeypatch, gpu_lib):
    """The column-sorted layout only reorders the gathers: every row still sums
    its products in row-major order, so SpMV results and (up to 3000 iterations
    of) whole solves equal the row-major sliced form bit for bit.  The FP32
    forms compare the FP32 row-major copy with the FP32 sorted layout (chunks of
    twice the entries)."""
    print(f"[case] sorted vs row-major slices {name} {precision}", flush=True)
    monkeypatch.setenv("HPRSOCP_FP32_OWN_LAYOUT", "0")
    a = _sliced_run(name, 8, False, monkeypatch, precision)
    monkeypatch.setenv("HPRSOCP_FP32_OWN_LAYOUT", "1")
    b = _sliced_run(name, 8, True, monkeypatch, precision)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])
    print(f"[result] {name} {precision}: iter {a[2].iter} / {b[2].iter} status {a[2].status} "
          f"kkt {a[2].residuals:.2e}", flush=True)
    assert np.isfinite(a[2].residuals) and np.all(np.isfinite(a[2].x)) and np.all(np.isfinite(a[2].y))
    assert a[2].iter == b[2].iter and a[2].primal_obj == b[2].primal_obj
    np.testing.assert_array_equal(a[2].x, b[2].x)
    np.testing.assert_array_equal(a[2].y, b[2].y)


@pytest.mark.parametrize("precision", ["fp32_refine", "fp32_delta"])
@pytest.mark.parametrize("name", ["c1_small", "qp_factored"])
def test_fp32_sorted_layouts_agree(name, precision, monkeypatch, gpu_lib):
    """An FP32 operator runs either its own sorted layout (16 K-entry chunks) or
    the FP64 layout (8 K) under the same float kernel, which takes the chunk
    size from the layout: both give the same iterates, bit for bit, and finite."""
    monkeypatch.setenv("HPRSOCP_FP32_OWN_LAYOUT", "0")
    a = _sliced_run(name, 8, True, monkeypatch, precision)
    monkeypatch.setenv("HPRSOCP_FP32_OWN_LAYOUT", "1")
    b = _sliced_run(name, 8, True, monkeypatch, precision)
    print(f"[result] {name} {precision}: iter {a[2].iter} / {b[2].iter} kkt {a[2].residuals:.2e} / "
          f"{b[2].residuals:.2e}", flush=True)
    assert np.isfinite(a[2].residuals) and np.all(np.isfinite(a[2].x)) and np.all(np.isfinite(a[2].y))
    assert a[2].iter == b[2].iter and a[2].primal_obj == b[2].primal_obj
    np.testing.assert_array_equal(a[2].x, b[2].x)
    np.testing.assert_array_equal(a[2].y, b[2].y)


def test_sorted_slices_multi_chunk_rows(monkeypatch, gpu_lib):
    """Rows with more entries in a slice than one sorted chunk holds (55000
    entries over 3 slices): SpMV vs scipy and bit-identical to row-major."""
    print("[case] sorted slices, rows longer than a chunk", flush=True)
    a = _sliced_run("dense_rows", 3, False, monkeypatch, solve=False)
    b = _sliced_run("dense_rows", 3, True, monkeypatch, solve=False)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])


@pytest.mark.parametrize("precision", ["fp32_refine", "fp32_delta"])
def test_sorted_slices_multi_chunk_rows_fp32(precision, monkeypatch, gpu_lib):
    """The same rows in the FP32 sorted layout (18 K entries per slice, more than
    one FP32 chunk): 200 iterations equal the row-major FP32 form bit for bit."""
    print(f"[case] FP32 sorted slices, rows longer than a chunk ({precision})", flush=True)
    monkeypatch.setenv("HPRSOCP_FP32_OWN_LAYOUT", "0")
    a = _sliced_run("dense_rows", 3, False, monkeypatch, precision, max_iter=200)
    monkeypatch.setenv("HPRSOCP_FP32_OWN_LAYOUT", "1")
    b = _sliced_run("dense_rows", 3, True, monkeypatch, precision, max_iter=200)
    assert np.all(np.isfinite(a[2].x)) and np.all(np.isfinite(a[2].y))
    assert a[2].iter == b[2].iter
    np.testing.assert_array_equal(a[2].x, b[2].x)
    np.testing.assert_array_equal(a[2].y, b[2].y)


@pytest.mark.parametrize("precision", ["fp64", "fp32_delta"])
@pytest.mark.parametrize("name", ["c1_small", "dense_rows"])
def test_sorted_kernel_block_sizes_agree(name, precision, monkeypatch, gpu_lib):
    """The sorted-slice kernel runs with 512 or 1024 threads per workgroup (the
    autotuner times both per operator): each row still sums its products in
    row-major order, so SpMVs and solves are the same bit for bit."""
    print(f"[case] sorted kernel 512 vs 1024 threads {name} {precision}", flush=True)
    solve = name != "dense_rows"
    runs = []
    for blk in ("512", "1024"):
        monkeypatch.setenv("HPRSOCP_SORT_BLOCK", blk)
        runs.append(_sliced_run(name, 8 if solve else 3, True, monkeypatch, precision, solve=solve, max_iter=1500))
    a, b = runs
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])
    if solve:
        print(f"[result] {name} {precision}: iter {a[2].iter} / {b[2].iter}", flush=True)
        assert a[2].iter == b[2].iter and a[2].primal_obj == b[2].primal_obj
        np.testing.assert_array_equal(a[2].x, b[2].x)
        np.testing.assert_array_equal(a[2].y, b[2].y)


@pytest.mark.parametrize("fused", ["1", "2"])
@pytest.mark.parametrize("precision", ["fp64", "fp32_refine", "fp32_delta"])
@pytest.mark.parametrize("name", ["c1_small", "c2_tiny", "qp_factored", "dense_rows"])
def test_slice_fused_bit_identical(name, precision, fused, monkeypatch, gpu_lib):
    """The slice-fused sorted kernel (one workgroup walks every slice of its row
    tile and writes finished row sums; tiles cut to 1 or 2 per workgroup slot)
    sums each slice's products in row-major order and the slices in slice order,
    as the unfused kernel and the tile kernel's partial sum do: SpMVs and whole
    solves equal the unfused sorted form bit for bit, with 512 and 1024 threads,
    and the FP32 copies (own layout or the FP64 one) agree the same way."""
    print(f"[case] slice-fused vs unfused sorted {name} {precision} tiles x{fused}", flush=True)
    solve = name != "dense_rows"
    P = 8 if solve else 3
    monkeypatch.setenv("HPRSOCP_FP32_OWN_LAYOUT", "1")
    monkeypatch.setenv("HPRSOCP_SLICE_FUSED", "0")
    a = _sliced_run(name, P, True, monkeypatch, precision, solve=solve, max_iter=1500)
    monkeypatch.setenv("HPRSOCP_SLICE_FUSED", fused)
    runs = []
    for blk, own in (("512", "1"), ("1024", "1"), ("512", "0")):
        monkeypatch.setenv("HPRSOCP_SORT_BLOCK", blk)
        monkeypatch.setenv("HPRSOCP_FP32_OWN_LAYOUT", own)
        runs.append(_sliced_run(name, P, True, monkeypatch, precision, solve=solve, max_iter=1500))
    for b in runs:
        np.testing.assert_array_equal(a[0], b[0])
        np.testing.assert_array_equal(a[1], b[1])
        if solve:
            assert np.all(np.isfinite(a[2].x)) and np.all(np.isfinite(a[2].y))
            assert a[2].iter == b[2].iter and a[2].primal_obj == b[2].primal_obj
            np.testing.assert_array_equal(a[2].x, b[2].x)
            np.testing.assert_array_equal(a[2].y, b[2].y)
    if solve:
        print(f"[result] {name} {precision}: iter {a[2].iter} status {a[2].status}", flush=True)


def test_slice_fused_form_reported(monkeypatch, gpu_lib):
    """A forced slice-fused layout shows in the plan report (form "sorted8-fused",
    FP32 copy "own-fused"), and the report is valid JSON with finite times."""
    monkeypatch.setenv("HPRSOCP_XCD_SLICES", "8")
    monkeypatch.setenv("HPRSOCP_XCD_SORTED", "1")
    monkeypatch.setenv("HPRSOCP_SLICE_FUSED", "1")
    with hprsocp.Solver(_slice_model("c1_small"), params(precision="fp32_delta")) as g:
        info = g.plan_info()
    ops = {o["name"]: o for o in info["operators"]}
    assert ops["A"]["form"] == "sorted8-fused" and ops["A^T"]["form"] == "sorted8-fused"
    assert ops["A"]["form32"] == "own-fused"
    assert all(np.isfinite(o["ms_form"]) for o in info["operators"])


@pytest.mark.parametrize("precision", ["fp64", "fp32_delta"])
@pytest.mark.parametrize("name", ["c1_small", "qp_factored", "dense_rows"])
def test_sorted_whole_rounds_bit_identical(name, precision, monkeypatch, gpu_lib):
    """The unfused sorted layout cut into whole rounds of workgroup slots (more,
    nnz-balanced row tiles) sums every row as the natural tiling does: SpMVs and
    solves are the same bit for bit."""
    print(f"[case] sorted slices in whole rounds vs natural tiles {name} {precision}", flush=True)
    solve = name != "dense_rows"
    P = 8 if solve else 3
    monkeypatch.setenv("HPRSOCP_SLICE_FUSED", "0")
    monkeypatch.setenv("HPRSOCP_SLICE_ROUNDS", "0")
    a = _sliced_run(name, P, True, monkeypatch, precision, solve=solve, max_iter=1500)
    monkeypatch.setenv("HPRSOCP_SLICE_ROUNDS", "1")
    b = _sliced_run(name, P, True, monkeypatch, precision, solve=solve, max_iter=1500)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])
    if solve:
        assert a[2].iter == b[2].iter and a[2].primal_obj == b[2].primal_obj
        np.testing.assert_array_equal(a[2].x, b[2].x)
        np.testing.assert_array_equal(a[2].y, b[2].y)


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "fp32_refine", "fp32_delta"])
def test_factored_q_primal_fusion_bit_identical(monkeypatch, prec):
    """With a factored Q and no variable cones the primal step runs in the F^T
    product's tile kernel; its iterates and solve equal the separate Q w product
    followed by the primal step bit for bit."""
    model = synthetic.make_c1(scale=0.02, seed=21, q_form="factored")
    outs, infos, res = [], [], []
    for fuse in ("0", "1"):
        monkeypatch.setenv("HPRSOCP_FUSE_QW", fuse)
        p = hprsocp.HPRSOCP_parameters(verbose=False, stoptol=1e-8, precision=prec)
        with hprsocp.Solver(model, p) as g:
            g.estimate_lambda()
            infos.append(g.plan_info()["qw_primal_fused"])
            for k in range(30):
                g.update(k, k % 10 == 9)
            outs.append({v: g.get(v).copy() for v in ("x", "y", "x_bar", "y_bar")})
        with hprsocp.Solver(model, p) as g:
            r = g.solve()
            res.append((r.status, r.iter, r.primal_obj, r.x.copy()))
    assert infos == [0, 1]
    for v in outs[0]:
        assert np.array_equal(outs[0][v], outs[1][v]), v
    assert res[0][:3] == res[1][:3]
    assert np.array_equal(res[0][3], res[1][3])


@pytest.mark.gpu
def test_factored_q_primal_fusion_without_diagonal_and_with_variable_cones(monkeypatch):
    """Q = F'F with no diagonal part: the fused primal step (no D w term) equals
    the separate kernels bit for bit.  With variable cones the steps stay apart."""
    import dataclasses
    base = synthetic.make_c1(scale=0.02, seed=22, q_form="factored")
    model = dataclasses.replace(base, Q=("factored", base.Q[1], None))
    outs = []
    for fuse in ("0", "1"):
        monkeypatch.setenv("HPRSOCP_FUSE_QW", fuse)
        p = hprsocp.HPRSOCP_parameters(verbose=False, stoptol=1e-6, precision="fp64")
        with hprsocp.Solver(model, p) as g:
            g.estimate_lambda()
            assert g.plan_info()["qw_primal_fused"] == int(fuse)
            for k in range(30):
                g.update(k, k % 10 == 9)
            outs.append({v: g.get(v).copy() for v in ("x", "y", "x_bar", "y_bar")})
    for v in outs[0]:
        assert np.array_equal(outs[0][v], outs[1][v]), v
    monkeypatch.delenv("HPRSOCP_FUSE_QW")
    cones = synthetic.make_small_qp(seed=5, q="factored")
    assert len(cones.SOC_var_idx) > 1
    with hprsocp.Solver(cones, hprsocp.HPRSOCP_parameters(verbose=False, stoptol=1e-6)) as g:
        assert g.plan_info()["qw_primal_fused"] == 0
