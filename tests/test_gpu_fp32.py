"""Mixed precision (precision "fp32_refine") against the FP64 oracle, through the C ABI.

In this mode the update kernels (SpMV, box/cone projections, Halpern averaging)
run on FP32 copies of the scaled operators and iterates; cone norms are
accumulated in FP64; residuals, M-norm, restarts and the sigma update run in
FP64 on the widened state at every check; FP64 iterations take over once the
KKT residual reaches refine_tol (default 1e-6) or stalls.

Tolerances (FP32 unit roundoff u = 2^-24 ~ 6e-8; entries of the scaled data are O(1)):
  * one HPR step from the same FP32-representable state:
        |a-b|_inf <= 2e-5 (1 + |b|_inf)       (a few roundings of sums of O(100) terms)
  * 30 HPR steps:  |a-b|_inf <= 1e-3 (1 + |b|_inf)   (the HPR map is nonexpansive, so the
        per-step rounding adds up at most linearly)
  * end-to-end at stoptol 1e-6: KKT (evaluated in FP64) < 1e-6, independent FP64
    certificate < 1e-5, primal objective within 1e-5 relative of an FP64 oracle solve
    at 1e-9 (the same bars as the FP64 path)
  * end-to-end at stoptol 1e-8 (FP64 iterations take over): KKT < 1e-8, certificate
    < 1e-7, objective within 1e-7 relative of the 1e-9 oracle solve.
"""
import numpy as np
import pytest

import hprsocp
from kat_models import params
from kkt import certificate
from test_gpu_parity import model

pytestmark = pytest.mark.gpu

F32 = dict(precision="fp32_refine")
STATE = {"noq": ["x", "y", "last_x", "last_y"],
         "qp": ["x", "y", "w", "ATy", "last_x", "last_y", "last_w", "last_ATy"]}
OUT = ["x_bar", "z_bar", "x_hat", "dx", "y_bar", "s", "dy"]
NAMES = ["c0", "noq_small", "qp_diag", "qp_csr", "qp_factored", "huge_cones", "huge_noq",
         "c1_tiny", "c2_tiny"]


def err(a, b):
    a, b = np.asarray(a), np.asarray(b)
    fin = np.isfinite(b)
    assert np.array_equal(np.isfinite(a), fin)
    if not fin.any():
        return 0.0
    return float(np.max(np.abs(a[fin] - b[fin])) / (1 + np.max(np.abs(b[fin]))))


def f32(v):
    return np.asarray(v, dtype=np.float32).astype(np.float64)


def start_pair(name, oracle_lib, warm=20):
    """GPU solver in FP32 mode and the oracle, both at the same FP32-representable
    state reached by `warm` oracle steps."""
    m = model(name)
    print(f"[case] fp32 {name} m={m.A.shape[0]} n={m.A.shape[1]} nnz={m.A.nnz}", flush=True)
    g = hprsocp.Solver(m, params(**F32))
    o = hprsocp.Solver(m, params(), lib=oracle_lib, prefix="oracle_")
    so = o.estimate_lambda()
    for s in (g, o):
        s.set_scalars(sigma=so["sigma"], lambda_A=so["lambda_max_A"], lambda_Q=so["lambda_max_Q"])
    kind = "qp" if so["has_quadratic"] else "noq"
    for k in range(warm):
        o.update(k, k == warm - 1)
    for v in STATE[kind] + (["w_bar"] if kind == "qp" else []):
        val = f32(o.get(v))
        o.set(v, val)
        g.set(v, val)
    return g, o, kind


@pytest.mark.parametrize("name", NAMES)
def test_fp32_one_step(name, oracle_lib, gpu_lib):
    g, o, kind = start_pair(name, oracle_lib)
    with g, o:
        g.update(20, True)
        o.update(20, True)
        worst = {v: err(g.get(v), o.get(v)) for v in STATE[kind] + OUT}
        print(f"[result] {name} one-step max rel err {max(worst.values()):.2e}", flush=True)
        for v, e in worst.items():
            assert e <= 2e-5, (v, e)


@pytest.mark.parametrize("name", NAMES)
def test_fp32_thirty_steps(name, oracle_lib, gpu_lib):
    g, o, kind = start_pair(name, oracle_lib)
    with g, o:
        for k in range(30):
            full = (k % 7 == 6) or k == 29
            g.update(20 + k, full)
            o.update(20 + k, full)
        worst = {v: err(g.get(v), o.get(v)) for v in STATE[kind] + OUT}
        print(f"[result] {name} 30-step max rel err {max(worst.values()):.2e}", flush=True)
        for v, e in worst.items():
            assert e <= 1e-3, (v, e)
        # FP64 residuals of the widened state agree with the oracle's on its own state
        rg, ro = g.residuals(1), o.residuals(1)
        assert rg["kkt"] == pytest.approx(ro["kkt"], rel=1e-2, abs=1e-6)


_ref9 = {}


def oracle_1e9(name, oracle_lib):
    if name not in _ref9:
        po = params(stoptol=1e-9, max_iter=400000, time_limit=300.0)
        with hprsocp.Solver(model(name), po, lib=oracle_lib, prefix="oracle_") as o:
            o.estimate_lambda()
            _ref9[name] = o.solve()
    return _ref9[name]


SOLVES = ["c0", "qp_diag", "qp_csr", "qp_factored", "huge_cones", "c1_tiny", "c2_tiny"]


@pytest.mark.parametrize("tol", [1e-6, 1e-8])
@pytest.mark.parametrize("name", SOLVES)
def test_fp32_refine_solve(name, tol, oracle_lib, gpu_lib):
    m = model(name)
    print(f"[case] fp32 solve {name} stoptol {tol:g}", flush=True)
    with hprsocp.Solver(m, params(stoptol=tol, max_iter=200000, time_limit=120.0, **F32)) as g:
        g.estimate_lambda()
        r = g.solve()
    ref = oracle_1e9(name, oracle_lib)
    rel = abs(r.primal_obj - ref.primal_obj) / (1 + abs(ref.primal_obj))
    cert = certificate(m, r.x, r.y, r.z)
    print(f"[result] {name} tol {tol:g}: {r.status} iter {r.iter} (fp32 {r.iter_fp32}) kkt {r.residuals:.2e} "
          f"cert Rp {cert['err_Rp']:.1e} Rd {cert['err_Rd']:.1e} obj rel err {rel:.2e}", flush=True)
    assert r.status == "OPTIMAL", (r.status, r.residuals, r.iter)
    assert r.residuals < tol
    assert 0 < r.iter_fp32 <= r.iter
    bar = 1e-5 if tol >= 1e-6 else 1e-7
    assert cert["err_Rp"] < bar and cert["err_Rd"] < bar, cert
    assert rel <= bar


def test_fp32_rejects_bad_params(gpu_lib):
    with pytest.raises(hprsocp.HPRSOCPError):
        hprsocp.Solver(model("c0"), params(precision="fp32_refine", fp32_stall_checks=0))
