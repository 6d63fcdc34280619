"""Precision "fp32_delta" against the FP64 oracle, through the C ABI.

In this mode the iterate is kept as the FP64 Halpern anchor of the restart
cycle (last_x, last_y, last_w, last_ATy) plus an FP32 difference; the SpMVs run
in FP32 on differences from the anchors, the raw values and projections in
FP64, and check iterations in FP64 on the joined state.  The rounding error of
a step is therefore a few FP32 units of the *difference* from the anchors, not
of the iterate: the tests measure the error against D = the largest distance
of the iterate from its anchors, which is far below the iterate's size.

Tolerances (FP32 unit roundoff u = 2^-24 ~ 6e-8):
  * one partial step from the same FP64 state, anchors 10 steps behind:
        |a-b|_inf <= 1e-5 D + 1e-12 (1 + |b|_inf)
    (a term dropped or doubled in the delta algebra would be O(|b|), far above)
  * 30 steps with FP64 check steps among them:
        |a-b|_inf <= 1e-4 D + 1e-11 (1 + |b|_inf)
  * end-to-end at stoptol 1e-6 and 1e-8: KKT < stoptol, independent FP64
    certificate and primal objective within the FP64 path's bars (1e-5 / 1e-7
    of an FP64 oracle solve at 1e-9), in about the FP64 path's iteration count.
"""
import numpy as np
import pytest

import hprsocp
from kat_models import params
from kkt import certificate
from test_gpu_fp32 import NAMES, SOLVES, oracle_1e9
from test_gpu_parity import model

pytestmark = pytest.mark.gpu

DELTA = dict(precision="fp32_delta")
STATE = {"noq": ["x", "y"], "qp": ["x", "y", "w", "ATy"]}
ANCHOR = {"x": "last_x", "y": "last_y", "w": "last_w", "ATy": "last_ATy"}


def abs_err(a, b):
    a, b = np.asarray(a), np.asarray(b)
    assert np.array_equal(np.isfinite(a), np.isfinite(b))
    fin = np.isfinite(b)
    return float(np.max(np.abs(a[fin] - b[fin]))) if fin.any() else 0.0


def inf(v):
    v = np.asarray(v)
    return float(np.max(np.abs(v))) if v.size else 0.0


def start_pair(name, oracle_lib, warm=200, lag=10):
    """GPU solver in delta mode and the oracle at the same FP64 state: `warm`
    oracle steps, anchors set to the state there (as a restart does), then
    `lag` more steps, so the iterate is `lag` steps away from its anchors."""
    m = model(name)
    print(f"[case] delta {name} m={m.A.shape[0]} n={m.A.shape[1]} nnz={m.A.nnz}", flush=True)
    g = hprsocp.Solver(m, params(**DELTA))
    o = hprsocp.Solver(m, params(), lib=oracle_lib, prefix="oracle_")
    so = o.estimate_lambda()
    for s in (g, o):
        s.set_scalars(sigma=so["sigma"], lambda_A=so["lambda_max_A"], lambda_Q=so["lambda_max_Q"])
    kind = "qp" if so["has_quadratic"] else "noq"
    for k in range(warm):
        o.update(k, False)
    for v in STATE[kind]:
        o.set(ANCHOR[v], o.get(v))
    for k in range(lag):
        o.update(k, False)
    for v in STATE[kind]:
        for name_ in (v, ANCHOR[v]):
            g.set(name_, o.get(name_))
    return g, o, kind


def distance(o, kind):
    return max(inf(np.asarray(o.get(v)) - np.asarray(o.get(ANCHOR[v]))) for v in STATE[kind])


@pytest.mark.parametrize("name", NAMES)
def test_delta_one_step(name, oracle_lib, gpu_lib):
    g, o, kind = start_pair(name, oracle_lib)
    with g, o:
        g.update(10, False)
        o.update(10, False)
        D = distance(o, kind)
        worst = {}
        for v in STATE[kind]:
            b = o.get(v)
            worst[v] = abs_err(g.get(v), b) / (1e-5 * D + 1e-12 * (1 + inf(b)))
        print(f"[result] {name} one step: D {D:.2e}, |x| {inf(o.get('x')):.2e}, "
              f"worst err / bar {max(worst.values()):.3f}", flush=True)
        for v, e in worst.items():
            assert e <= 1.0, (v, e)


@pytest.mark.parametrize("name", NAMES)
def test_delta_thirty_steps(name, oracle_lib, gpu_lib):
    g, o, kind = start_pair(name, oracle_lib)
    with g, o:
        for k in range(30):
            full = (k % 7 == 6) or k == 29
            g.update(10 + k, full)
            o.update(10 + k, full)
        D = distance(o, kind)
        outs = ["x_bar", "z_bar", "dx", "y_bar", "s", "dy"]
        worst = {}
        for v in STATE[kind] + outs:
            b = o.get(v)
            worst[v] = abs_err(g.get(v), b) / (1e-4 * D + 1e-11 * (1 + inf(b)))
        print(f"[result] {name} 30 steps: D {D:.2e}, worst err / bar {max(worst.values()):.3f}", flush=True)
        for v, e in worst.items():
            assert e <= 1.0, (v, e)


_ref64 = {}


def gpu_fp64(name, tol):
    key = (name, tol)
    if key not in _ref64:
        with hprsocp.Solver(model(name), params(stoptol=tol, max_iter=200000, time_limit=120.0)) as g:
            g.estimate_lambda()
            _ref64[key] = g.solve()
    return _ref64[key]


@pytest.mark.parametrize("tol", [1e-6, 1e-8])
@pytest.mark.parametrize("name", SOLVES)
def test_delta_solve(name, tol, oracle_lib, gpu_lib):
    m = model(name)
    print(f"[case] delta solve {name} stoptol {tol:g}", flush=True)
    with hprsocp.Solver(m, params(stoptol=tol, max_iter=200000, time_limit=120.0, **DELTA)) as g:
        g.estimate_lambda()
        r = g.solve()
    ref = oracle_1e9(name, oracle_lib)
    r64 = gpu_fp64(name, tol)
    rel = abs(r.primal_obj - ref.primal_obj) / (1 + abs(ref.primal_obj))
    cert = certificate(m, r.x, r.y, r.z)
    print(f"[result] {name} tol {tol:g}: {r.status} iter {r.iter} (fp32 {r.iter_fp32}; fp64 path {r64.iter}) "
          f"kkt {r.residuals:.2e} cert Rp {cert['err_Rp']:.1e} Rd {cert['err_Rd']:.1e} obj rel err {rel:.2e}",
          flush=True)
    assert r.status == "OPTIMAL", (r.status, r.residuals, r.iter)
    assert r.residuals < tol
    assert 0 < r.iter_fp32 < r.iter
    bar = 1e-5 if tol >= 1e-6 else 1e-7
    assert cert["err_Rp"] < bar and cert["err_Rd"] < bar, cert
    assert rel <= bar
    # the FP64 path's trajectory, up to rounding-sized shifts of restart decisions
    assert r.iter <= 1.25 * r64.iter + 2 * 150


def test_delta_graph_matches_eager(gpu_lib):
    """Captured check windows replay the same kernels: bit-identical iterates."""
    m = model("qp_factored")
    out = []
    for graph in (True, False):
        with hprsocp.Solver(m, params(stoptol=1e-12, max_iter=1000, use_graph=graph, **DELTA)) as g:
            g.estimate_lambda()
            r = g.solve()
            out.append((r.iter, r.x.copy(), r.y.copy()))
    assert out[0][0] == out[1][0] == 1000
    np.testing.assert_array_equal(out[0][1], out[1][1])
    np.testing.assert_array_equal(out[0][2], out[1][2])


@pytest.mark.parametrize("name", ["qp_factored", "c0"])
def test_delta_timing_leaves_the_state(name, oracle_lib, gpu_lib):
    """hprsocp_time_iterations runs partial steps and then joins the deltas
    into the FP64 state: the anchors are untouched and the iterate is the
    oracle's after the same steps (the FP32 form's mirror promotion must not
    run on the delta form's mirrors, which hold differences)."""
    g, o, kind = start_pair(name, oracle_lib)
    with g, o:
        anchors = {ANCHOR[v]: np.asarray(o.get(ANCHOR[v])).copy() for v in STATE[kind]}
        g.time_iterations(5)
        for k in range(5):
            o.update(k, False)
        D = distance(o, kind)
        for a, val in anchors.items():
            np.testing.assert_array_equal(g.get(a), val)
        for v in STATE[kind]:
            b = o.get(v)
            assert abs_err(g.get(v), b) <= 1e-4 * D + 1e-11 * (1 + inf(b)), v
