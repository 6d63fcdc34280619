"""Batched solves (hprsocp_batch_*): many problems packed block-diagonally and
iterated in lockstep, each instance with its own scaling, sigma, restarts and
termination.

Checks, per instance:
  * a batch of one is the single-problem solve (same tiles, same trajectory):
    identical iteration count and objective to the last bit;
  * in a mixed-size batch every instance reaches KKT < stoptol = 1e-6, its
    independent FP64 certificate is < 1e-5, and its objective agrees with the
    single-problem GPU solve to 1e-6 relative and with an FP64 oracle solve at
    1e-9 to 1e-5 relative (the bars of the single-problem tests).  Iteration
    counts equal the single solve's up to summation-order effects (tiles of a
    batch can pick other SpMV lane widths), bounded here by 25%;
  * warm starts (per-instance initial points) and FP32 mode go through the same
    per-instance machinery.
"""
import numpy as np
import pytest

import hprsocp
from hprsocp import synthetic
from kat_models import params
from kkt import certificate

pytestmark = pytest.mark.gpu


def noq_models():
    out = [synthetic.make_c0(seed=s) for s in (0, 1)]
    out += [synthetic.make_small_qp(seed=20 + k, n1=40 + 30 * k, var_sizes=(3, 5, 8 + k), m1=10 + 5 * k,
                                    con_sizes=(4, 3 + k), density=0.2, q="none") for k in range(4)]
    return out


def diag_models():
    return [synthetic.make_small_qp(seed=40 + k, n1=30 + 25 * k, var_sizes=(3, 4 + k), m1=8 + 4 * k,
                                    con_sizes=(3, 5), density=0.25, q="diag") for k in range(4)]


def single_solve(model, **kw):
    with hprsocp.Solver(model, params(**kw)) as g:
        g.estimate_lambda()
        return g.solve()


def batch_solve(models, **kw):
    with hprsocp.BatchSolver(models, params(**kw)) as b:
        b.estimate_lambda()
        return b.solve()


def test_batch_of_one_is_the_single_solve(gpu_lib):
    m = synthetic.make_c0(seed=0)
    print("[case] batch of one (c0)", flush=True)
    r1 = single_solve(m, stoptol=1e-6)
    rb = batch_solve([m], stoptol=1e-6)[0]
    assert rb.status == r1.status == "OPTIMAL"
    assert rb.iter == r1.iter
    assert rb.primal_obj == r1.primal_obj
    np.testing.assert_array_equal(rb.x, r1.x)
    np.testing.assert_array_equal(rb.y, r1.y)


_o9 = {}


def oracle9(model, key, oracle_lib):
    if key not in _o9:
        with hprsocp.Solver(model, params(stoptol=1e-9, max_iter=400000, time_limit=300.0),
                            lib=oracle_lib, prefix="oracle_") as o:
            o.estimate_lambda()
            _o9[key] = o.solve()
    return _o9[key]


@pytest.mark.parametrize("kind", ["noq", "diag"])
def test_batch_matches_single_solves(kind, oracle_lib, gpu_lib):
    models = noq_models() if kind == "noq" else diag_models()
    print(f"[case] batch {kind} of {len(models)}", flush=True)
    rb = batch_solve(models, stoptol=1e-6, max_iter=100000)
    for i, (m, r) in enumerate(zip(models, rb)):
        rs = single_solve(m, stoptol=1e-6, max_iter=100000)
        ro = oracle9(m, (kind, i), oracle_lib)
        cert = certificate(m, r.x, r.y, r.z)
        rel = abs(r.primal_obj - rs.primal_obj) / (1 + abs(rs.primal_obj))
        rel9 = abs(r.primal_obj - ro.primal_obj) / (1 + abs(ro.primal_obj))
        print(f"[result] {kind}[{i}] m={m.A.shape[0]} n={m.A.shape[1]}: batch iter {r.iter} single {rs.iter} "
              f"kkt {r.residuals:.2e} obj rel {rel:.2e} vs oracle1e-9 {rel9:.2e}", flush=True)
        assert r.status == "OPTIMAL", (i, r.status, r.residuals)
        assert r.residuals < 1e-6
        assert cert["err_Rp"] < 1e-5 and cert["err_Rd"] < 1e-5, (i, cert)
        assert rel <= 1e-6
        assert rel9 <= 1e-5
        assert 0.8 * rs.iter <= r.iter <= 1.25 * rs.iter + 150


@pytest.mark.parametrize("precision", ["fp64", "fp32_delta"])
def test_batch_warm_start(precision, gpu_lib):
    """Half the instances warm-started (the C4 setting).  With fp32_delta the
    step schedule of an instance depends on the batch's restarts, so a cold
    instance follows its cold-batch trajectory only to rounding."""
    models = noq_models()[:4]
    cold = batch_solve(models, stoptol=1e-6, precision=precision)
    rng = np.random.default_rng(5)
    xs = [r.x * (1 + 0.01 * rng.standard_normal(r.x.shape)) if i % 2 == 0 else None for i, r in enumerate(cold)]
    ys = [r.y * (1 + 0.01 * rng.standard_normal(r.y.shape)) if i % 2 == 0 else None for i, r in enumerate(cold)]
    with hprsocp.BatchSolver(models, params(stoptol=1e-6, precision=precision)) as b:
        b.estimate_lambda()
        b.set_initial(xs, ys)
        warm = b.solve()
    for i, (c, w) in enumerate(zip(cold, warm)):
        print(f"[result] warm[{i}] cold iter {c.iter} warm iter {w.iter} obj {w.primal_obj:.10e}", flush=True)
        assert w.status == "OPTIMAL"
        assert abs(w.primal_obj - c.primal_obj) <= 1e-5 * (1 + abs(c.primal_obj))
        if i % 2 == 1:  # cold instances of the warm batch follow the cold trajectory
            if precision == "fp64":
                assert w.iter == c.iter
            else:
                assert abs(w.iter - c.iter) <= 300


def test_batch_fp32(gpu_lib):
    models = noq_models()
    rb = batch_solve(models, stoptol=1e-6, precision="fp32_refine")
    for i, (m, r) in enumerate(zip(models, rb)):
        cert = certificate(m, r.x, r.y, r.z)
        print(f"[result] fp32 batch[{i}] iter {r.iter} (fp32 {r.iter_fp32}) kkt {r.residuals:.2e}", flush=True)
        assert r.status == "OPTIMAL"
        assert cert["err_Rp"] < 1e-5 and cert["err_Rd"] < 1e-5


@pytest.mark.parametrize("precision", ["fp32_refine", "fp32_delta"])
@pytest.mark.parametrize("kind", ["noq", "diag"])
def test_batch_fp32_instances_follow_single_solves(kind, precision, gpu_lib):
    """fp32_refine: each instance switches from FP32 to FP64 on its own
    (refine_tol or stall), so an instance of a batch follows its
    single-problem solve bit for bit: same iterations, same FP32 iterations,
    same objective.  fp32_delta: each instance's anchors, images and deltas
    are its own, but the batch runs an FP64 step (instead of an FP32 delta
    step) after any instance's restart, so the step schedule, and with it the
    rounding, depends on the batch: the instance follows its single solve to
    rounding (same iteration count up to two check windows, objective to 1e-7)."""
    models = noq_models() if kind == "noq" else diag_models()
    print(f"[case] {precision} batch {kind} vs single solves", flush=True)
    rb = batch_solve(models, stoptol=1e-8, precision=precision)
    for i, (m, r) in enumerate(zip(models, rb)):
        rs = single_solve(m, stoptol=1e-8, precision=precision)
        print(f"[result] {kind}[{i}] batch iter {r.iter} (fp32 {r.iter_fp32}) single {rs.iter} (fp32 {rs.iter_fp32})",
              flush=True)
        assert r.status == rs.status == "OPTIMAL"
        if precision == "fp32_delta":
            assert abs(r.iter - rs.iter) <= 300 and 0 < r.iter_fp32 < r.iter
            assert abs(r.primal_obj - rs.primal_obj) <= 1e-7 * (1 + abs(rs.primal_obj))
            continue
        assert (r.iter, r.iter_fp32) == (rs.iter, rs.iter_fp32)
        assert r.primal_obj == rs.primal_obj
        np.testing.assert_array_equal(r.x, rs.x)


def test_batch_rejects_mixed_q(gpu_lib):
    a = synthetic.make_small_qp(seed=1, q="none")
    b = synthetic.make_small_qp(seed=2, q="diag")
    with pytest.raises(hprsocp.HPRSOCPError):
        hprsocp.BatchSolver([a, b], params())
    c = synthetic.make_small_qp(seed=3, q="factored")
    with pytest.raises(hprsocp.HPRSOCPError):
        hprsocp.BatchSolver([c], params())


def test_portfolio_contiguous_rows(oracle_lib, gpu_lib):
    """Portfolio instances (the C3 family, small): every row of A is one run of
    consecutive columns (the budget row and the dense F^T rows), which the
    lane-group SpMV reads without column indices (CsrT::c0).  SpMV against
    scipy, batch = single solves bit for bit, and objectives against the
    oracle's solve."""
    import scipy.sparse as sp
    models = synthetic.make_c3(count=4, n_lo=200, n_hi=3000)
    for m in models:
        A = m.A.tocsr()
        for r in range(A.shape[0]):
            idx = A.indices[A.indptr[r]:A.indptr[r + 1]]
            assert np.array_equal(idx, np.arange(idx[0], idx[0] + len(idx)) if len(idx) else idx)
    m = models[0]
    with hprsocp.Solver(m, params()) as g:
        A = m.A.tocsr()
        As = sp.csr_matrix((g.scaled_A(), A.indices, A.indptr), shape=A.shape)
        rng = np.random.default_rng(3)
        v, w = rng.standard_normal(A.shape[1]), rng.standard_normal(A.shape[0])
        np.testing.assert_allclose(g.spmv(v), As @ v, rtol=1e-12, atol=1e-12 * np.abs(As @ v).max())
        np.testing.assert_allclose(g.spmv(w, trans=True), As.T @ w, rtol=1e-12,
                                   atol=1e-12 * np.abs(As.T @ w).max())
    rb = batch_solve(models, stoptol=1e-6)
    for i, (m, r) in enumerate(zip(models, rb)):
        rs = single_solve(m, stoptol=1e-6)
        with hprsocp.Solver(m, params(stoptol=1e-6), lib=oracle_lib, prefix="oracle_") as o:
            o.estimate_lambda()
            ro = o.solve()
        print(f"[result] portfolio[{i}] n={m.A.shape[1]}: batch iter {r.iter} single {rs.iter} oracle {ro.iter} "
              f"obj {r.primal_obj:.10e} / {ro.primal_obj:.10e}", flush=True)
        assert r.status == rs.status == ro.status == "OPTIMAL"
        assert r.iter == rs.iter and r.primal_obj == rs.primal_obj
        assert abs(r.primal_obj - ro.primal_obj) <= 1e-6 * (1 + abs(ro.primal_obj))


@pytest.mark.parametrize("precision", ["fp64", "fp32_refine", "fp32_delta"])
def test_batch_general_q(precision, oracle_lib, gpu_lib):
    """Batches of QPs with a general (non-diagonal) CSR Q, whose Q applies run as
    tile SpMVs over the block-diagonal Q: every instance is optimal, its
    objective matches the oracle's single solve, and retirements (uneven sizes)
    leave the others running."""
    models = [synthetic.make_small_qp(seed=60 + k, n1=30 + 20 * k, var_sizes=(3, 4 + k), m1=8 + 4 * k,
                                      con_sizes=(3, 5), density=0.25, q="csr") for k in range(4)]
    print(f"[case] general-Q batch ({precision})", flush=True)
    rb = batch_solve(models, stoptol=1e-8, precision=precision)
    for i, (m, r) in enumerate(zip(models, rb)):
        with hprsocp.Solver(m, params(stoptol=1e-9, max_iter=200000), lib=oracle_lib, prefix="oracle_") as o:
            o.estimate_lambda()
            ro = o.solve()
        rel = abs(r.primal_obj - ro.primal_obj) / (1 + abs(ro.primal_obj))
        cert = certificate(m, r.x, r.y, r.z)
        print(f"[result] general-Q[{i}] {precision}: iter {r.iter} obj rel err {rel:.2e} "
              f"cert Rp {cert['err_Rp']:.1e} Rd {cert['err_Rd']:.1e}", flush=True)
        assert r.status == "OPTIMAL" and ro.status == "OPTIMAL"
        assert rel <= 1e-7
        assert cert["err_Rp"] < 1e-7 and cert["err_Rd"] < 1e-7
