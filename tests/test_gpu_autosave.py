"""auto_save: the checkpoint of save_state_to_hdf5! (reference
src/algorithm.jl:306-423) written at every print step and at termination,
with the best-iterate record of compute_residuals! (algorithm.jl:176-195).

The first test mirrors the reference's own auto-save test
(test/runtests.jl:895-984: the same keys, the auto_save flag, best residuals
below the tolerance, best/iteration <= current/iteration, vector lengths).
The others pin what that test leaves open:
  * the last checkpoint of a solve is the returned solution, bit for bit, with
    the returned residuals, iteration and sigma;
  * best/x_org, y_org, z_org are the iterate whose residuals best/ records: an
    independent FP64 certificate of those vectors reproduces best/err_Rp,
    err_Rd and primal_obj, and best/KKTx_and_gap is their maximum;
  * the best record is the running minimum of the KKT value over the residual
    steps: with print_frequency = check_iter every residual step is logged,
    and best/ holds the logged minimum (first occurrence);
  * a solve warm-started from the best iterate of an interrupted solve
    (resume_parameters) reaches the cold solve's optimum in fewer iterations;
  * a write failure leaves the solve running; batches reject auto_save.
"""
import re

import numpy as np
import pytest

import hprsocp
from kat_models import params
from kkt import certificate
from test_gpu_parity import model

pytestmark = pytest.mark.gpu


def run(m, path, **kw):
    with hprsocp.Solver(m, params(auto_save=True, save_filename=str(path), **kw)) as g:
        g.estimate_lambda()
        return g.solve()


def test_autosave_reference_checks(tmp_path, gpu_lib):
    """runtests.jl:895-984 on a constrained QP (check_iter 100, print every 500, stoptol 1e-4)."""
    m = model("qp_csr")
    path = tmp_path / "test_autosave_runtests.npz"
    print(f"[case] autosave reference checks -> {path}", flush=True)
    r = run(m, path, max_iter=5000, stoptol=1e-4, time_limit=120.0, check_iter=100, print_frequency=500)
    assert r.status == "OPTIMAL" and r.residuals < 1e-4
    assert path.exists()
    d = hprsocp.load_autosave(str(path))
    for k in ("iteration", "x_org", "w_org", "z_org", "y_org", "sigma", "err_Rp", "err_Rd", "primal_obj",
              "dual_obj", "rel_gap"):
        assert "current/" + k in d, k
    for k in ("iteration", "x_org", "w_org", "z_org", "y_org", "sigma", "err_Rp", "err_Rd"):
        assert "best/" + k in d, k
    assert "parameters/stoptol" in d
    assert d["parameters/auto_save"] is True
    xb, wb, zb, yb = d["best/x_org"], d["best/w_org"], d["best/z_org"], d["best/y_org"]
    assert len(xb) > 0 and len(wb) > 0 and len(zb) > 0 and len(yb) > 0
    assert len(xb) == len(wb) == len(zb) == m.A.shape[1] and len(yb) == m.A.shape[0]
    assert d["best/err_Rp"] < 1e-4 and d["best/err_Rd"] < 1e-4
    assert 0 <= d["best/iteration"] <= d["current/iteration"]
    # the remaining groups of algorithm.jl:352-418
    for k in ("time_elapsed", "timestamp", "KKTx_and_gap"):
        assert "current/" + k in d
    for k in ("primal_obj", "dual_obj", "rel_gap", "KKTx_and_gap"):
        assert "best/" + k in d
    p = params(stoptol=1e-4)
    assert d["parameters/stoptol"] == 1e-4 and d["parameters/check_iter"] == 100
    assert d["parameters/print_frequency"] == 500 and d["parameters/max_iter"] == 5000
    assert d["parameters/sigma"] == p.sigma and d["parameters/ruiz_iterations"] == p.ruiz_iterations
    assert d["parameters/bc_scaling_norm_type"] == "l2"
    assert d["parameters/soc_block_scaling_strategy"] == "phase_taper"
    assert d["parameters/warm_up"] is False and d["parameters/use_Ruiz_scaling"] is True
    assert re.fullmatch(r"\d{4}-\d\d-\d\dT\d\d:\d\d:\d\d\.\d{3}", d["current/timestamp"])
    assert "parameters/initial_x" not in d


@pytest.mark.parametrize("name,precision", [("qp_csr", "fp64"), ("c0", "fp64"), ("qp_csr", "fp32_refine"),
                                            ("qp_csr", "fp32_delta"), ("c0", "fp32_delta")])
def test_autosave_last_checkpoint_is_the_result(name, precision, tmp_path, gpu_lib):
    m = model(name)
    path = tmp_path / "state.npz"
    print(f"[case] autosave final state {name} {precision}", flush=True)
    r = run(m, path, stoptol=1e-6, precision=precision)
    d = hprsocp.load_autosave(str(path))
    assert d["current/iteration"] == r.iter
    np.testing.assert_array_equal(d["current/x_org"], r.x)
    np.testing.assert_array_equal(d["current/y_org"], r.y)
    np.testing.assert_array_equal(d["current/z_org"], r.z)
    np.testing.assert_array_equal(d["current/w_org"], r.w)
    assert d["current/err_Rp"] == r.err_Rp and d["current/err_Rd"] == r.err_Rd
    assert d["current/primal_obj"] == r.primal_obj and d["current/dual_obj"] == r.dual_obj
    assert d["current/rel_gap"] == r.gap and d["current/KKTx_and_gap"] == r.residuals
    assert d["current/sigma"] == r.sigma
    # the final iterate has KKT < stoptol: the best record is at least as good
    assert d["best/KKTx_and_gap"] <= d["current/KKTx_and_gap"]
    if d["best/iteration"] == d["current/iteration"]:
        np.testing.assert_array_equal(d["best/x_org"], r.x)
    if name == "c0":  # no quadratic term: w is x (collect_results!)
        np.testing.assert_array_equal(d["current/w_org"], d["current/x_org"])
        np.testing.assert_array_equal(d["best/w_org"], d["best/x_org"])


def _check_best_is_its_iterate(m, d):
    cert = certificate(m, d["best/x_org"], d["best/y_org"], d["best/z_org"])
    print(f"[result] best iter {d['best/iteration']} err_Rp {d['best/err_Rp']:.6e} / cert {cert['err_Rp']:.6e}, "
          f"err_Rd {d['best/err_Rd']:.6e} / cert {cert['err_Rd']:.6e}", flush=True)
    assert d["best/KKTx_and_gap"] == max(d["best/err_Rp"], d["best/err_Rd"], d["best/rel_gap"])
    np.testing.assert_allclose(cert["err_Rp"], d["best/err_Rp"], rtol=1e-6, atol=1e-13)
    np.testing.assert_allclose(cert["err_Rd"], d["best/err_Rd"], rtol=1e-6, atol=1e-13)
    np.testing.assert_allclose(cert["primal_obj"], d["best/primal_obj"], rtol=1e-10, atol=1e-12)


_LOG = re.compile(r"^\s*(\d+)\s+(\S+)\s+(\S+)\s+(\S+)\s+(\S+)\s+(\S+)\s+(\S+)\s+(\S+)\s*$")


def test_autosave_best_is_running_minimum(tmp_path, capfd, gpu_lib):
    """Interrupted solve (MAX_ITER): every residual step is logged; best/ is the
    logged minimum of max(errRp, errRd, gap) and its vectors certify its residuals."""
    m = model("qp_diag")
    path = tmp_path / "interrupted.npz"
    print("[case] autosave running minimum (qp_diag, max_iter 600)", flush=True)
    capfd.readouterr()
    r = run(m, path, stoptol=1e-12, max_iter=600, check_iter=50, print_frequency=50, verbose=True)
    out = capfd.readouterr().out
    assert r.status == "MAX_ITER"
    log = {}
    for line in out.splitlines():
        g = _LOG.match(line)
        if g:
            it = int(g.group(1))
            rp, rd, gap = float(g.group(2)), float(g.group(3)), float(g.group(6))
            log[it] = max(rp, rd, gap)
    assert sorted(log) == list(range(0, 601, 50)), sorted(log)
    d = hprsocp.load_autosave(str(path))
    assert d["current/iteration"] == 600
    kmin = min(log.values())
    # the log rounds to 3 significant digits (0.5% at most): best/ must be a
    # logged minimum up to twice that rounding
    assert abs(d["best/KKTx_and_gap"] - kmin) <= 0.011 * kmin
    near = [it for it, v in log.items() if v <= kmin * 1.011]
    assert d["best/iteration"] in near
    _check_best_is_its_iterate(m, d)


def test_autosave_resume_from_best(tmp_path, gpu_lib):
    m = model("qp_csr")
    cold = run(m, tmp_path / "cold.npz", stoptol=1e-6)
    path = tmp_path / "partial.npz"
    part = run(m, path, stoptol=1e-6, max_iter=max(cold.iter // 2, 150))
    assert part.status == "MAX_ITER"
    d = hprsocp.load_autosave(str(path))
    _check_best_is_its_iterate(m, d)
    p = hprsocp.resume_parameters(str(path), params(stoptol=1e-6))
    np.testing.assert_array_equal(p.initial_x, d["best/x_org"])
    assert p.sigma == d["best/sigma"]
    warm = hprsocp.optimize(m, p)
    print(f"[result] resume: cold {cold.iter} iterations, interrupted at {part.iter} "
          f"(best KKT {d['best/KKTx_and_gap']:.2e} at {d['best/iteration']}), warm {warm.iter}", flush=True)
    assert warm.status == "OPTIMAL"
    assert abs(warm.primal_obj - cold.primal_obj) <= 1e-5 * (1 + abs(cold.primal_obj))
    assert warm.iter < cold.iter
    # the initial points given in the parameters are part of the checkpoint
    p2 = hprsocp.resume_parameters(str(path), params(stoptol=1e-6, auto_save=True,
                                                     save_filename=str(tmp_path / "warm.npz")))
    hprsocp.optimize(m, p2)
    d2 = hprsocp.load_autosave(str(tmp_path / "warm.npz"))
    np.testing.assert_array_equal(d2["parameters/initial_x"], d["best/x_org"])
    np.testing.assert_array_equal(d2["parameters/initial_y"], d["best/y_org"])


def test_autosave_write_failure_keeps_solving(tmp_path, gpu_lib):
    m = model("c0")
    bad = tmp_path / "missing_dir" / "state.npz"
    r = run(m, bad, stoptol=1e-6)
    assert r.status == "OPTIMAL"
    assert not bad.exists()


def test_autosave_rejected_for_batches(gpu_lib):
    with pytest.raises(hprsocp.HPRSOCPError):
        hprsocp.BatchSolver([model("c0")], params(auto_save=True))


def test_autosave_without_rows(tmp_path, gpu_lib):
    """m = 0 (the reference's "SOC variables only" solve, runtests.jl:544-562):
    no y_org datasets, as save_state_to_hdf5! writes them only when m > 0."""
    from kat_models import kats, model_from
    s = kats()["solve_soc_variables_only"]
    m = model_from(s)
    path = tmp_path / "norows.npz"
    r = run(m, path, stoptol=s["stoptol"], max_iter=s["max_iter"])
    assert r.status == "OPTIMAL"
    d = hprsocp.load_autosave(str(path))
    assert "current/y_org" not in d and "best/y_org" not in d
    assert d["best/x_org"].shape == (3,)
    np.testing.assert_array_equal(d["current/x_org"], r.x)
    p = hprsocp.resume_parameters(str(path))
    assert p.initial_y is None and p.initial_x.shape == (3,)
