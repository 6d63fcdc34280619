"""Behaviour of optimize() as the reference's API tests check it
(test/runtests.jl):

* parameter settings: box models solve to OPTIMAL at stoptol 1e-3 and 1e-5
  (runtests.jl:836-857), with the box optimum x = l pinned here as well;
* result fields: x, y, z, w, iter, time, residuals, primal_obj, gap, status
  (runtests.jl:859-890);
* verbose = false prints nothing, also from restarts and sigma updates
  (runtests.jl:489-527 checks check_restart and update_sigma! one call at a
  time; here a whole solve that restarts and updates sigma is captured);
* an unavailable device_number (runtests.jl:1050-1071).  The reference falls
  back to its CPU path there; this library has no CPU path, so the request
  fails with an error that names the device.
"""
import dataclasses
import re

import numpy as np
import pytest
import scipy.sparse as sp

import hprsocp
from test_gpu_parity import model

pytestmark = pytest.mark.gpu


def box_model(n, c=None, l=None, u=None):
    """make_box_socp_model of runtests.jl:59-68: min c'x over l <= x <= u, no rows."""
    c = np.ones(n) if c is None else np.asarray(c, float)
    l = np.zeros(n) if l is None else np.asarray(l, float)
    u = np.ones(n) if u is None else np.asarray(u, float)
    return hprsocp.build_from_SOCP_data(sp.csr_matrix((n, n)), c, sp.csr_matrix((0, n)), np.zeros(0), [1], 0, 0,
                                        l, u, [n + 1], verbose=False)


def quiet(**kw):
    base = dict(max_iter=50000, time_limit=600.0, warm_up=False, verbose=False, print_frequency=-1)
    base.update(kw)
    return hprsocp.HPRSOCP_parameters(**base)


@pytest.mark.parametrize("tol", [1e-3, 1e-5])
def test_parameter_settings(tol, gpu_lib):
    print(f"[case] box model n=5, stoptol {tol}", flush=True)
    r = hprsocp.optimize(box_model(5), quiet(stoptol=tol))
    assert r.status == "OPTIMAL" and r.residuals < tol
    np.testing.assert_allclose(r.x, np.zeros(5), atol=10 * tol)
    assert abs(r.primal_obj) <= 10 * tol


def test_result_fields(gpu_lib):
    print("[case] box model n=4, result fields", flush=True)
    r = hprsocp.optimize(box_model(4), quiet(stoptol=1e-6))
    names = {f.name for f in dataclasses.fields(r)}
    for k in ("x", "y", "z", "w", "iter", "time", "residuals", "primal_obj", "gap", "status"):
        assert k in names, k
    assert r.status == "OPTIMAL"
    assert r.x.shape == (4,) and r.z.shape == (4,) and r.w.shape == (4,) and r.y.shape == (0,)
    assert r.iter > 0 and r.time >= 0.0 and r.residuals < 1e-6


def test_verbose_false_is_silent(capfd, gpu_lib):
    """A solve that restarts and updates sigma prints nothing with verbose=false,
    and prints its log with verbose=true (so the capture itself works)."""
    m = model("qp_csr")
    capfd.readouterr()
    r = hprsocp.optimize(m, quiet(stoptol=1e-8, warm_up=True, print_frequency=100))
    out = capfd.readouterr()
    assert r.status == "OPTIMAL" and r.restarts > 0
    assert out.out == "" and out.err == "", (out.out[:400], out.err[:400])
    r2 = hprsocp.optimize(m, quiet(stoptol=1e-8, verbose=True, print_frequency=100))
    out = capfd.readouterr()
    print(f"[result] silent solve: {r.iter} iterations, {r.restarts} restarts; verbose log "
          f"{len(out.out.splitlines())} lines", flush=True)
    assert r2.iter == r.iter and len(out.out.splitlines()) > 5


def test_unavailable_device_number(gpu_lib):
    with pytest.raises(hprsocp.HPRSOCPError, match="device_number 999"):
        hprsocp.optimize(box_model(10, u=10 * np.ones(10)), quiet(device_number=999, max_iter=100))
    # the default device solves the same model
    r = hprsocp.optimize(box_model(10, u=10 * np.ones(10)), quiet(max_iter=100))
    assert r.status in ("OPTIMAL", "MAX_ITER")


_LOG = re.compile(r"^\s*(\d+)\s+(\S+)\s+(\S+)\s+(\S+)\s+(\S+)\s+(\S+)\s+(\S+)\s+(\S+)\s*$")
_MILESTONE = re.compile(r"^KKT < 1e-(4|6) at iter = (\d+)$")


def test_kkt_milestones(capfd, gpu_lib):
    """iter_4 / iter_6 (update_milestone_tracking!, algorithm.jl:2693-2718): the
    first residual step whose KKT value is below 1e-4 / 1e-6.  With
    print_frequency = check_iter every residual step is logged, so the
    milestones can be read off the log (3 significant digits, hence the 1.1%
    margin); an unreached milestone reports the final iteration and time
    (algorithm.jl:1763-1767)."""
    m = model("qp_csr")
    capfd.readouterr()
    r = hprsocp.optimize(m, quiet(stoptol=1e-8, check_iter=50, print_frequency=50, verbose=True))
    out = capfd.readouterr().out
    assert r.status == "OPTIMAL"
    log, said = {}, {}
    for line in out.splitlines():
        g = _LOG.match(line)
        if g:
            log[int(g.group(1))] = max(float(g.group(2)), float(g.group(3)), abs(float(g.group(6))))
        g = _MILESTONE.match(line.strip())
        if g:
            said[int(g.group(1))] = int(g.group(2))
    print(f"[result] iter_4 {r.iter_4} iter_6 {r.iter_6} iter {r.iter}; log {len(log)} residual steps", flush=True)
    assert said == {4: r.iter_4, 6: r.iter_6}
    assert 0 < r.iter_4 <= r.iter_6 <= r.iter and 0 < r.time_4 <= r.time_6 <= r.time
    for tol, it in ((1e-4, r.iter_4), (1e-6, r.iter_6)):
        assert it in log and log[it] < tol * 1.011
        assert all(v >= tol / 1.011 for k, v in log.items() if k < it)
    # a solve that ends before 1e-6: iter_6 and time_6 are the final ones
    r5 = hprsocp.optimize(m, quiet(stoptol=1e-5))
    assert r5.status == "OPTIMAL" and r5.iter_4 <= r5.iter and r5.iter_6 == r5.iter
    if r5.residuals >= 1e-6:
        assert r5.time_6 == r5.time
