"""Captured check windows (params.use_graph): runs of partial HPR updates are
captured once into a HIP graph and replayed; every solve below replays its
graphs many times, with the eager checks, restarts and host synchronisations
between replays.  The graph path must give the eager path's results bit for
bit (same kernels, same order, same device-side scalars)."""
import numpy as np
import pytest

import hprsocp
from kat_models import params
from test_gpu_batch import noq_models, diag_models
from test_gpu_parity import model

pytestmark = pytest.mark.gpu


def solve(m, **kw):
    with hprsocp.Solver(m, params(**kw)) as g:
        g.estimate_lambda()
        return g.solve()


@pytest.mark.parametrize("name", ["c0", "qp_diag", "qp_csr", "qp_factored", "huge_cones", "c2_tiny"])
@pytest.mark.parametrize("precision", ["fp64", "fp32_refine"])
def test_graph_equals_eager(name, precision, gpu_lib):
    m = model(name)
    print(f"[case] graph vs eager {name} {precision}", flush=True)
    rg = solve(m, stoptol=1e-6, use_graph=True, precision=precision)
    re = solve(m, stoptol=1e-6, use_graph=False, precision=precision)
    print(f"[result] {name} {precision}: iter {rg.iter} / {re.iter}", flush=True)
    assert rg.status == re.status == "OPTIMAL"
    assert rg.iter == re.iter
    assert rg.primal_obj == re.primal_obj
    np.testing.assert_array_equal(rg.x, re.x)
    np.testing.assert_array_equal(rg.y, re.y)


@pytest.mark.parametrize("kind", ["noq", "diag"])
def test_batch_graph_equals_eager(kind, gpu_lib):
    models = noq_models() if kind == "noq" else diag_models()
    out = []
    for ug in (True, False):
        with hprsocp.BatchSolver(models, params(stoptol=1e-6, use_graph=ug)) as b:
            b.estimate_lambda()
            out.append(b.solve())
    for i, (a, e) in enumerate(zip(*out)):
        assert a.iter == e.iter, i
        assert a.primal_obj == e.primal_obj, i
        np.testing.assert_array_equal(a.x, e.x)
