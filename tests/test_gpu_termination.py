"""Termination and host-side options of the solve loop (algorithm.jl:2765-2905,
check_break algorithm.jl:1169-1187, optimize's warm-up algorithm.jl:2538-2581):

* a time limit ends the solve with TIME_LIMIT and a finite iterate;
* print steps only add residual evaluations: the iterate after a given number
  of iterations does not depend on print_frequency;
* optimize's 200-iteration warm-up solve does not change the main solve.
"""
import numpy as np
import pytest

import hprsocp
from kat_models import params
from test_gpu_parity import model

pytestmark = pytest.mark.gpu


def solve(m, **kw):
    with hprsocp.Solver(m, params(**kw)) as g:
        g.estimate_lambda()
        return g.solve()


@pytest.mark.parametrize("precision", ["fp64", "fp32_delta"])
def test_time_limit(precision, gpu_lib):
    m = model("c1_tiny")
    print(f"[case] time limit 0.01 s on c1_tiny ({precision})", flush=True)
    r = solve(m, stoptol=1e-14, time_limit=0.01, max_iter=10**7, precision=precision)
    print(f"[result] status {r.status} after {r.iter} iterations, {r.time:.3f} s", flush=True)
    assert r.status == "TIME_LIMIT"
    assert r.iter > 0 and np.isfinite(r.residuals)
    assert np.all(np.isfinite(r.x)) and np.all(np.isfinite(r.y))


@pytest.mark.parametrize("freq", [37, 500])
@pytest.mark.parametrize("name", ["qp_csr", "c0"])
def test_print_frequency_leaves_the_iterates(name, freq, gpu_lib):
    """Print steps evaluate residuals (and so may end a solve at a different
    residual step), but they must not perturb the iteration: run to the same
    max_iter, both solves hold the same iterate bit for bit."""
    m = model(name)
    print(f"[case] print_frequency {freq} vs adaptive on {name}", flush=True)
    a = solve(m, stoptol=1e-15, max_iter=1200, print_frequency=-1)
    b = solve(m, stoptol=1e-15, max_iter=1200, print_frequency=freq)
    assert a.status == b.status == "MAX_ITER"
    assert a.iter == b.iter == 1200 and a.primal_obj == b.primal_obj and a.restarts == b.restarts
    np.testing.assert_array_equal(a.x, b.x)
    np.testing.assert_array_equal(a.y, b.y)


def test_warm_up_leaves_the_solve(gpu_lib):
    m = model("qp_diag")
    print("[case] optimize with and without warm-up", flush=True)
    a = hprsocp.optimize(m, params(stoptol=1e-6, warm_up=False))
    b = hprsocp.optimize(m, params(stoptol=1e-6, warm_up=True))
    assert a.status == b.status == "OPTIMAL"
    assert a.iter == b.iter and a.primal_obj == b.primal_obj
    np.testing.assert_array_equal(a.x, b.x)
