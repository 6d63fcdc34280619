"""The robust-regression batch (the mixed-size QP-SOCP batch with half the
instances warm-started), reduced to 4 instances with d in [2e3, 1e4].

Each instance's size-(d+1) constraint cone spans many 256-row tiles, so its
projection is the two-phase multi-tile path, and instances retire while other
instances' multi-tile cones keep running (compacted plans, HUGE-cone slots).
For FP64 and the FP32 delta form, at stoptol 1e-6 and 1e-8, every instance
  * terminates OPTIMAL with KKT < stoptol;
  * passes the independent FP64 certificate (unscaled data) at 10 x stoptol;
  * matches the oracle's solve from the same warm start at the same stoptol:
    FP64 in iteration count (same algorithm, same decisions) and objective to
    1e-9 relative; FP32 delta objective to stoptol relative.
"""
import numpy as np
import pytest

import hprsocp
from hprsocp import synthetic
from kat_models import params
from kkt import certificate

pytestmark = pytest.mark.gpu

_models = {}
_oracle = {}


def c4_small():
    if "m" not in _models:
        _models["m"] = synthetic.make_c4(count=4, seed=4000, d_lo=2000, d_hi=10000)
    return _models["m"]


def oracle_solve(oracle_lib, i, tol):
    key = (i, tol)
    if key not in _oracle:
        models, warms = c4_small()
        with hprsocp.Solver(models[i], params(stoptol=tol, max_iter=400000, time_limit=300.0),
                            lib=oracle_lib, prefix="oracle_") as o:
            o.estimate_lambda()
            if warms[i] is not None:
                o.set_initial(warms[i], None)
            _oracle[key] = o.solve()
    return _oracle[key]


@pytest.mark.parametrize("tol", [1e-6, 1e-8])
@pytest.mark.parametrize("precision", ["fp64", "fp32_delta"])
def test_c4_batch(precision, tol, oracle_lib, gpu_lib):
    models, warms = c4_small()
    print(f"[case] c4 batch of {len(models)} ({precision}, stoptol {tol:g}); d+1 = "
          f"{[m.A.shape[0] - m.number_linear_con for m in models]}", flush=True)
    with hprsocp.BatchSolver(models, params(stoptol=tol, max_iter=400000, precision=precision)) as b:
        b.estimate_lambda()
        b.set_initial(warms, None)
        rb = b.solve()
    for i, (m, r) in enumerate(zip(models, rb)):
        assert m.A.shape[0] - m.number_linear_con > 4 * 256  # a multi-tile cone
        ro = oracle_solve(oracle_lib, i, tol)
        cert = certificate(m, r.x, r.y, r.z)
        rel = abs(r.primal_obj - ro.primal_obj) / (1 + abs(ro.primal_obj))
        print(f"[result] c4[{i}] {'warm' if warms[i] is not None else 'cold'} n={m.A.shape[1]}: "
              f"iter {r.iter} (fp32 {r.iter_fp32}) oracle {ro.iter}; kkt {r.residuals:.2e}; obj rel {rel:.2e}; "
              f"cert Rp {cert['err_Rp']:.1e} Rd {cert['err_Rd']:.1e} gap {cert['rel_gap']:.1e}", flush=True)
        assert r.status == ro.status == "OPTIMAL", (i, r.status, r.residuals)
        assert r.residuals < tol
        assert cert["err_Rp"] < 10 * tol and cert["err_Rd"] < 10 * tol, (i, cert)
        assert cert["y_cone_violation"] < 10 * tol * (1 + np.abs(r.y).max())
        if precision == "fp64":
            assert r.iter == ro.iter
            assert rel <= 1e-9
        else:
            assert 0 < r.iter_fp32 < r.iter
            assert rel <= tol
