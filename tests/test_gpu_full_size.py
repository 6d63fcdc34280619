"""The four benchmark workloads at their full sizes, checked through properties
that hold at any size (the oracle is too slow to solve them in a test):

  * C1, the large single QP-SOCP (n = 1e6, nnz(A) = 3.0e7, factored Q): the
    sorted 8-slice A and A^T products the tuner picks at this size agree with
    scipy's on the scaled matrix to 1e-13 relative; the FP64 solve terminates
    OPTIMAL at stoptol 1e-6 and its returned (x, y, z) passes the independent
    FP64 KKT certificate (unscaled data) at 1e-5, the reference's relative
    measures (src/algorithm.jl:77-175) recomputed outside the solver;
  * C2, the cone-projection stress QP-SOCP (1e5 size-3 + 100 size-10,000 cones
    per side): the stand-alone projection on that cone layout agrees with an
    exactly rounded projection to 1e-14 relative to each cone's scale; FP64 and
    FP32 delta solves reach KKT < 1e-6 with the certificate at 1e-5;
  * C3, the 512 portfolio SOCPs, and C4, the 64 robust-regression QP-SOCPs with
    half of them warm-started: every instance of the FP64 batch terminates
    OPTIMAL with KKT < 1e-6 and the certificate at 1e-5, and a sample of
    instances (the smallest, the largest and others across the size range)
    matches its single-problem GPU solve: objective to 1e-6 relative and
    iteration count within 25 % (tiles of a batch may pick other SpMV lane
    widths, so summation orders differ).
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp

import hprsocp
from hprsocp import synthetic
from kat_models import params
from kkt import certificate
from test_gpu_parity import close

pytestmark = pytest.mark.gpu

try:  # loaded before the library's first HIP call, so both share torch's HIP runtime
    import torch
except ImportError:
    torch = None

TOL = 1e-6


def check_solution(m, r, tol=TOL, what="", verbose=True):
    cert = certificate(m, r.x, r.y, r.z)
    if verbose:
        print(f"[result] {what}: {r.status} iter {r.iter} kkt {r.residuals:.2e} time {r.time:.2f} s; "
              f"cert Rp {cert['err_Rp']:.1e} Rd {cert['err_Rd']:.1e} gap {cert['rel_gap']:.1e}", flush=True)
    assert r.status == "OPTIMAL", (what, r.status, r.residuals, r.iter)
    assert r.residuals < tol
    assert cert["err_Rp"] < 10 * tol and cert["err_Rd"] < 10 * tol, (what, cert)
    assert cert["y_cone_violation"] < 10 * tol * (1 + np.abs(r.y).max()), (what, cert)
    assert cert["z_cone_violation"] < 10 * tol * (1 + np.abs(r.z).max()), (what, cert)
    return cert


def test_c1_full_size(gpu_lib):
    m = synthetic.make_c1(scale=1.0, seed=1000, q_form="factored")
    print(f"[case] c1 full size: m={m.A.shape[0]} n={m.A.shape[1]} nnz(A)={m.A.nnz}", flush=True)
    with hprsocp.Solver(m, params(stoptol=TOL, max_iter=200000, time_limit=110.0)) as g:
        ops = {o["name"]: o for o in g.plan_info()["operators"]}
        print(f"[case] plan A {ops['A']['form']}, A^T {ops['A^T']['form']}", flush=True)
        A = m.A.tocsr()
        As = sp.csr_matrix((g.scaled_A(), A.indices, A.indptr), shape=A.shape)
        rng = np.random.default_rng(7)
        v, w = rng.standard_normal(A.shape[1]), rng.standard_normal(A.shape[0])
        close(g.spmv(v), As @ v, 1e-13)
        close(g.spmv(w, trans=True), As.T @ w, 1e-13)
        del As
        g.estimate_lambda()
        r = g.solve()
    check_solution(m, r, what="c1 fp64")


def exact_project(v, idx):
    """Pi_K of every cone [idx[k], idx[k+1]) with |xi| from an exactly rounded
    sum of squares; rows outside the cones are copied."""
    out = v.copy()
    sizes = np.diff(idx)
    for size in np.unique(sizes):
        starts = idx[:-1][sizes == size]
        seg = v[starts[:, None] + np.arange(size)]
        t, xi = seg[:, 0], seg[:, 1:]
        if size <= 3:
            ns = np.sqrt(np.array([math.fsum(r) for r in xi * xi]))
        else:
            ns = np.array([math.sqrt(math.fsum(r * r)) for r in xi])
        pt = 0.5 * (ns + t)
        proj = np.where((ns <= t)[:, None], seg,
                        np.where((ns <= -t)[:, None], 0.0,
                                 np.concatenate([pt[:, None], (pt / np.where(ns > 0, ns, 1))[:, None] * xi], 1)))
        out[starts[:, None] + np.arange(size)] = proj
    return out


def test_c2_full_size_cones(gpu_lib):
    if torch is None:
        pytest.skip("torch provides the device buffers here")
    m = synthetic.make_c2(scale=1.0, seed=2000)
    idx = np.asarray(m.SOC_var_idx, dtype=np.int64) - 1
    n = m.A.shape[1]
    print(f"[case] c2 full-size cone layout: {len(idx) - 1} cones over {n} rows", flush=True)
    rng = np.random.default_rng(3)
    v = rng.standard_normal(n)
    for k in range(0, len(idx) - 1, 3):  # a third of the cones inside, a third in the polar cone
        s, e = idx[k], idx[k + 1]
        nx = np.linalg.norm(v[s + 1:e])
        v[s] = 1.5 * nx + 0.1
        if k + 1 < len(idx) - 1:
            s1, e1 = idx[k + 1], idx[k + 2]
            v[s1] = -1.5 * np.linalg.norm(v[s1 + 1:e1]) - 0.1
    dv = torch.from_numpy(v).cuda()
    dout = torch.full_like(dv, 7.0)
    didx = torch.from_numpy(idx.astype(np.int32)).cuda()
    hprsocp.project_cones(dv.data_ptr(), dout.data_ptr(), n, didx.data_ptr(), len(idx) - 1)
    got = dout.cpu().numpy()
    want = exact_project(v, idx)
    scale = np.ones(n)
    for size in np.unique(np.diff(idx)):
        starts = idx[:-1][np.diff(idx) == size]
        rows = starts[:, None] + np.arange(size)
        scale[rows] = 1 + np.abs(v[rows]).max(1, keepdims=True)
    inside = np.zeros(n, bool)
    inside[idx[0]:idx[-1]] = True
    err = np.abs(got[inside] - want[inside]) / scale[inside]
    print(f"[result] c2 projection: max relative error {err.max():.2e}", flush=True)
    assert err.max() <= 1e-14
    np.testing.assert_array_equal(got[~inside], 7.0)  # rows outside every cone are not written


@pytest.mark.parametrize("precision", ["fp64", "fp32_delta"])
def test_c2_full_size_solve(precision, gpu_lib):
    m = synthetic.make_c2(scale=1.0, seed=2000)
    print(f"[case] c2 full size ({precision}): m={m.A.shape[0]} n={m.A.shape[1]} nnz(A)={m.A.nnz}", flush=True)
    with hprsocp.Solver(m, params(stoptol=TOL, max_iter=200000, time_limit=60.0, precision=precision)) as g:
        g.estimate_lambda()
        r = g.solve()
    check_solution(m, r, what=f"c2 {precision}")


def sample(sizes, k):
    order = np.argsort(sizes, kind="stable")
    picks = {int(order[0]), int(order[-1])}
    picks |= {int(order[j]) for j in np.linspace(0, len(order) - 1, k).round().astype(int)}
    return sorted(picks)


def batch_against_singles(models, warms, picks, what):
    with hprsocp.BatchSolver(models, params(stoptol=TOL, max_iter=400000, time_limit=100.0)) as b:
        b.estimate_lambda()
        if warms is not None:
            b.set_initial(warms, None)
        rb = b.solve()
    worst = {}
    for i, (m, r) in enumerate(zip(models, rb)):
        cert = check_solution(m, r, what=f"{what}[{i}]", verbose=i in picks)
        for k in ("err_Rp", "err_Rd", "rel_gap"):
            worst[k] = max(worst.get(k, 0.0), cert[k])
    print(f"[result] {what}: all {len(models)} OPTIMAL; worst certificate {worst}", flush=True)
    for i in picks:
        with hprsocp.Solver(models[i], params(stoptol=TOL, max_iter=400000, time_limit=60.0)) as g:
            g.estimate_lambda()
            if warms is not None and warms[i] is not None:
                g.set_initial(warms[i], None)
            r1 = g.solve()
        rel = abs(rb[i].primal_obj - r1.primal_obj) / (1 + abs(r1.primal_obj))
        print(f"[result] {what}[{i}] n={models[i].A.shape[1]}: batch iter {rb[i].iter}, single iter {r1.iter}, "
              f"obj rel {rel:.1e}", flush=True)
        assert r1.status == "OPTIMAL"
        assert rel <= TOL
        assert abs(rb[i].iter - r1.iter) <= 0.25 * r1.iter + 10, (i, rb[i].iter, r1.iter)


def test_c3_full_batch(gpu_lib):
    models = synthetic.make_c3(count=512, seed=3000)
    sizes = [m.A.shape[1] for m in models]
    picks = sample(sizes, 8)
    print(f"[case] c3 full batch of {len(models)} (n {min(sizes)}..{max(sizes)}); singles {picks}", flush=True)
    batch_against_singles(models, None, picks, "c3")


def test_c4_full_batch(gpu_lib):
    models, warms = synthetic.make_c4(count=64, seed=4000)
    sizes = [m.A.shape[1] for m in models]
    picks = sample(sizes, 4)
    print(f"[case] c4 full batch of {len(models)} (n {min(sizes)}..{max(sizes)}); singles {picks} "
          f"(warm: {[i for i in picks if warms[i] is not None]})", flush=True)
    batch_against_singles(models, warms, picks, "c4")
