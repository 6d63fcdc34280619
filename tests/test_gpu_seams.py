"""The fine-grained seams of the solve loop on the GPU, against the oracle's.

* hprsocp_update_sigma (update_sigma!, algorithm.jl:729-990) and
  hprsocp_restart (do_restart!, algorithm.jl:1130-1166) on identical state:
  sigma to 1e-12 (1e-7 through a golden-section search), restart fields
  exactly, restarted iterates to 1e-12;
* hprsocp_iterate (the loop of algorithm.jl:2806-2904 in windows): windows of
  any length give the iterates, iteration count, restart count and sigma of one
  hprsocp_solve bit for bit, and the same trajectory as the oracle's solve;
* a host loop that drives only the seams (update, residuals, M-norm,
  check_restart, update_sigma, restart) reproduces hprsocp_solve's trajectory.
"""
import ctypes as C
import math

import numpy as np
import pytest

import hprsocp
from hprsocp import _abi, synthetic
from kat_models import params

pytestmark = pytest.mark.gpu

MODELS = {
    "noq": lambda: synthetic.make_small_qp(seed=3, q="none"),
    "diag": lambda: synthetic.make_small_qp(seed=7, q="diag"),
    "csr": lambda: synthetic.make_small_qp(seed=8, q="csr"),
    "factored": lambda: synthetic.make_small_qp(seed=9, q="factored"),
    "c0": lambda: synthetic.make_c0(seed=0),
}
_cache = {}


def model(name):
    if name not in _cache:
        _cache[name] = MODELS[name]()
    return _cache[name]


STATE = ["x_bar", "last_x", "y_bar", "last_y", "w_bar", "last_w", "ATy_bar", "last_ATy"]


def _pair(name, oracle_lib, **kw):
    m = model(name)
    print(f"[case] seams {name} m={m.A.shape[0]} n={m.A.shape[1]}", flush=True)
    p = params(**kw)
    return hprsocp.Solver(m, p), hprsocp.Solver(m, p, lib=oracle_lib, prefix="oracle_")


def _same_state(g, o, rng):
    for v in STATE:
        a = rng.normal(size=g.m if v in _abi.VEC_IS_M else g.n)
        g.set(v, a)
        o.set(v, a)


@pytest.mark.parametrize("name", ["noq", "diag", "csr", "factored"])
@pytest.mark.parametrize("res", [dict(kkt=1.0, err_Rp=1e-3, err_Rd=1e-3),
                                 dict(kkt=2e-3, err_Rp=1e-3, err_Rd=1e-10),
                                 dict(kkt=1.5e-3, err_Rp=3e-8, err_Rd=4e-9)])
def test_update_sigma_seam(name, res, oracle_lib, gpu_lib):
    g, o = _pair(name, oracle_lib)
    with g, o:
        so = o.estimate_lambda()
        g.set_scalars(sigma=0.9, lambda_A=so["lambda_max_A"], lambda_Q=so["lambda_max_Q"])
        o.set_scalars(sigma=0.9, lambda_A=so["lambda_max_A"], lambda_Q=so["lambda_max_Q"])
        _same_state(g, o, np.random.default_rng(5))
        fields = dict(first_restart=0, restart_flag=1, best_kkt=1e-3, best_gap=1.0, current_gap=1.3,
                      best_sigma=2.0, best_iter=3)
        rg, ro = g.init_restart(), o.init_restart()
        for k, v in fields.items():
            setattr(rg, k, v)
            setattr(ro, k, v)
        g.update_sigma(rg, res)
        o.update_sigma(ro, res)
        sg, so2 = g.scalars()["sigma"], o.scalars()["sigma"]
        # QP proposals come from a golden-section search: with dots that differ in
        # summation order, comparisons of nearly equal model values near the
        # minimum can flip, so the minimiser is fixed only to ~sqrt(eps) relative
        tol = 1e-12 if name == "noq" else 1e-7
        print(f"[result] {name}: sigma gpu {sg:.15e} oracle {so2:.15e}", flush=True)
        assert sg == pytest.approx(so2, rel=tol)
        assert sg != 0.9
        assert rg.as_dict() == ro.as_dict()


@pytest.mark.parametrize("name", ["noq", "diag", "csr", "factored"])
def test_restart_seam(name, oracle_lib, gpu_lib):
    g, o = _pair(name, oracle_lib)
    with g, o:
        so = o.estimate_lambda()
        for s in (g, o):
            s.set_scalars(sigma=so["sigma"], lambda_A=so["lambda_max_A"], lambda_Q=so["lambda_max_Q"])
        _same_state(g, o, np.random.default_rng(6))
        for flag in (0, 2):
            rg, ro = g.init_restart(), o.init_restart()
            for r in (rg, ro):
                r.restart_flag, r.current_gap, r.inner, r.times, r.first_restart = flag, 0.25, 77, 4, 0
            g.restart(rg)
            o.restart(ro)
            assert rg.as_dict() == ro.as_dict()
            if flag:
                assert (rg.inner, rg.times, rg.last_gap) == (0, 5, 0.25) and math.isinf(rg.save_gap)
            for v in ["x", "last_x", "y", "last_y", "w", "last_w", "ATy", "last_ATy", "ATy_bar"]:
                a, b = g.get(v), o.get(v)
                np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-12 * (1 + np.abs(b).max()), err_msg=v)


@pytest.mark.parametrize("name,chunks", [("c0", [1, 149, 150, 151, 1000]), ("diag", [37]),
                                         ("factored", [300])])
@pytest.mark.parametrize("precision", ["fp64", "fp32_delta"])
def test_iterate_windows_match_solve(name, chunks, precision, gpu_lib):
    m = model(name)
    print(f"[case] iterate {name} {precision} chunks={chunks}", flush=True)
    p = params(stoptol=1e-6, max_iter=100000, precision=precision)
    with hprsocp.Solver(m, p) as g:
        g.estimate_lambda()
        x0, y0 = g.get("x"), g.get("y")
        full = g.solve()
    with hprsocp.Solver(m, p) as g:
        g.estimate_lambda()
        g.set("x", x0)
        g.set("y", y0)
        k = 0
        r = g.iterate(chunks[0], reset=True)
        calls = 1
        while r.status == "CONTINUE":
            k += 1
            assert r.iter == sum(chunks[min(j, len(chunks) - 1)] for j in range(k))
            r = g.iterate(chunks[min(k, len(chunks) - 1)])
            calls += 1
            assert calls < 100000
    print(f"[result] {name}: solve iter {full.iter} restarts {full.restarts}; iterate calls {calls} "
          f"iter {r.iter}", flush=True)
    assert r.status == full.status == "OPTIMAL"
    assert (r.iter, r.restarts, r.iter_fp32) == (full.iter, full.restarts, full.iter_fp32)
    assert r.sigma == full.sigma and r.primal_obj == full.primal_obj
    np.testing.assert_array_equal(r.x, full.x)
    np.testing.assert_array_equal(r.y, full.y)


def test_iterate_continue_result_and_reset(gpu_lib):
    m = model("c0")
    with hprsocp.Solver(m, params(stoptol=1e-6, max_iter=100000)) as g:
        g.estimate_lambda()
        r = g.iterate(300, reset=True)
        assert r.status == "CONTINUE" and r.iter == 300 and r.residuals > 0
        r2 = g.iterate(0)
        assert r2.status == "CONTINUE" and r2.iter == 300
        with pytest.raises(hprsocp.HPRSOCPError):
            g.iterate(-1)


@pytest.mark.parametrize("name", ["c0", "noq", "diag", "csr", "factored"])
def test_seam_loop_matches_solve(name, oracle_lib, gpu_lib):
    """The reference's loop (algorithm.jl:2806-2904) written on the host over
    the seams: compute_residuals!, check_break, check_restart, update_sigma!,
    do_restart!, the update, compute_M_norm!.  Same iterations and restarts as
    hprsocp_solve, sigma and objective to rounding."""
    m = model(name)
    print(f"[case] seam loop {name}", flush=True)
    p = params(stoptol=1e-6, max_iter=20000)
    with hprsocp.Solver(m, p) as g:
        g.estimate_lambda()
        ref = g.solve()
    check_iter = 150
    with hprsocp.Solver(m, p) as g:
        g.estimate_lambda()
        r = g.init_restart()
        r.best_sigma = g.scalars()["sigma"]
        res = None
        it = 0
        def print_step(t):  # print_step, algorithm.jl:1910-1912
            return max(10.0 ** math.floor(math.log10(t)) / 10.0, 10.0) if t > 0 else 10.0

        def print_yes(t):  # should_print with print_frequency = -1 (algorithm.jl:2810)
            return math.fmod(t, print_step(t)) == 0.0

        while True:
            if it % check_iter == 0 or print_yes(it):
                res = g.residuals(it)
                if res["kkt"] < p.stoptol:
                    break
            g.check_restart(r, it, check_iter, g.scalars()["sigma"], res["kkt"])
            g.update_sigma(r, res)
            g.restart(r)
            nx = it + 1
            to_check = nx % check_iter == 0 or r.restart_flag > 0 or print_yes(nx)
            g.update(r.inner, to_check)
            if r.restart_flag > 0 or nx % check_iter == 0:
                mn = g.mnorm()
                if r.restart_flag > 0:
                    r.last_gap = min(mn, r.last_gap)
                if nx % check_iter == 0:
                    r.current_gap = mn
            r.inner += 1
            it += 1
            assert it <= 20000
        sig = g.scalars()["sigma"]
    print(f"[result] {name}: seam loop iter {it} restarts {r.times} sigma {sig:.12e} | solve iter {ref.iter} "
          f"restarts {ref.restarts} sigma {ref.sigma:.12e}", flush=True)
    assert it == ref.iter and r.times == ref.restarts
    assert sig == pytest.approx(ref.sigma, rel=1e-12)
    assert res["primal_obj"] == pytest.approx(ref.primal_obj, rel=1e-9, abs=1e-12)
