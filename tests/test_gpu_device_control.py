"""Device control of batch solves: the window-edge step of the solve loop
(residual block, deferred M-norms, termination test, check_restart,
update_sigma!, do_restart! counters, milestones, retirement; algorithm.jl:2806-2867)
runs in a kernel queued behind each window, and the host only reads a status
vector one event late.  HPRSOCP_DEVICE_CONTROL=0 runs the same solve with
those steps on the host, from the same control source.

Checks:
  * the control arithmetic (own exp / log, golden section, sigma update)
    gives the same bits on the device and on the host;
  * in FP64 every instance of a device-controlled batch follows its
    host-controlled trajectory bit for bit: status, iteration, milestone
    iterations, restarts, sigma, residuals and solution arrays are equal, for
    LP, sparse-Q QP, portfolio, diagonal-Q QP and robust-regression batches
    (the diagonal-Q sigma update runs its golden-section walks on the device
    between two controller phases), at a KKT stop, at max_iter (a print step,
    not a check) and at a time limit;
  * in the FP32 delta form (where the device loop adds an FP64 step after
    every check) instances agree to rounding: iteration counts within two check
    windows, objectives to 1e-7 relative.
"""
import numpy as np
import pytest

import hprsocp
from hprsocp import synthetic
from kat_models import params

pytestmark = pytest.mark.gpu


def lp_models():
    out = [synthetic.make_c0(seed=s) for s in (0, 1)]
    out += [synthetic.make_small_qp(seed=20 + k, n1=40 + 30 * k, var_sizes=(3, 5, 8 + k), m1=10 + 5 * k,
                                    con_sizes=(4, 3 + k), density=0.2, q="none") for k in range(4)]
    return out


def csr_models():
    return [synthetic.make_small_qp(seed=60 + k, n1=30 + 20 * k, var_sizes=(3, 4 + k), m1=8 + 4 * k,
                                    con_sizes=(3, 5), density=0.25, q="csr") for k in range(4)]


def portfolio_models():
    return synthetic.make_c3(count=5, n_lo=200, n_hi=2500)


def diag_models():
    return [synthetic.make_small_qp(seed=40 + k, n1=30 + 25 * k, var_sizes=(3, 4 + k), m1=8 + 4 * k,
                                    con_sizes=(3, 5), density=0.25, q="diag") for k in range(4)]


def regression_models():
    """Robust regressions (diagonal Q, one SOC of d+1 rows split over many
    tiles), d in [2000, 6000], cold starts."""
    return synthetic.make_c4(count=4, seed=4100, d_lo=2000, d_hi=6000)[0]


MODELS = {"lp": lp_models, "csr": csr_models, "portfolio": portfolio_models, "diag": diag_models,
          "regression": regression_models}


def batch_solve(models, device, monkeypatch, warm=None, **kw):
    monkeypatch.setenv("HPRSOCP_DEVICE_CONTROL", "1" if device else "0")
    with hprsocp.BatchSolver(models, params(**kw)) as b:
        b.estimate_lambda()
        if warm is not None:
            b.set_initial(*warm)
        return b.solve()


def assert_same(rd, rh, tag):
    for i, (d, h) in enumerate(zip(rd, rh)):
        print(f"[result] {tag}[{i}] device iter {d.iter} ({d.status}, {d.restarts} restarts) "
              f"host iter {h.iter} ({h.status}, {h.restarts} restarts) kkt {d.residuals:.3e}", flush=True)
        assert (d.status, d.iter, d.iter_4, d.iter_6, d.restarts) == (h.status, h.iter, h.iter_4, h.iter_6,
                                                                         h.restarts), i
        assert d.sigma == h.sigma, i
        assert (d.residuals, d.primal_obj, d.dual_obj, d.err_Rp, d.err_Rd, d.gap) == \
            (h.residuals, h.primal_obj, h.dual_obj, h.err_Rp, h.err_Rd, h.gap), i
        np.testing.assert_array_equal(d.x, h.x)
        np.testing.assert_array_equal(d.y, h.y)
        np.testing.assert_array_equal(d.z, h.z)


def test_control_arithmetic_is_the_same_on_device_and_host(gpu_lib):
    rng = np.random.default_rng(11)
    xs = np.concatenate([
        [0.0, -0.0, 1.0, -1.0, 0.5, 2.0, 1e-300, 1e-12, 1e12, 700.0, -700.0, 709.7, -745.0, 1e-320],
        rng.standard_normal(300) * 10.0,
        np.exp(rng.uniform(-60, 60, 300)),
        -np.exp(rng.uniform(-30, 30, 100)),
    ])
    m = synthetic.make_c0(seed=0)
    with hprsocp.Solver(m, params()) as g:
        dev, host = g.probe_ctl(xs)
    assert dev.shape == (xs.size, 5)
    np.testing.assert_array_equal(dev, host)
    # the own exp / log are accurate to a few ulp against the C library
    fin = np.abs(xs) < 700
    np.testing.assert_allclose(host[fin, 0], np.exp(xs[fin]), rtol=4e-16, atol=0)
    pos = np.abs(xs) > 0
    np.testing.assert_allclose(host[pos, 1], np.log(np.abs(xs[pos])), rtol=4e-16, atol=1e-300)
    np.testing.assert_array_equal(host[:, 2], np.sqrt(np.abs(xs)))


@pytest.mark.parametrize("kind", ["lp", "csr", "portfolio", "diag", "regression"])
def test_fp64_device_control_is_the_host_trajectory(kind, monkeypatch, gpu_lib):
    models = MODELS[kind]()
    print(f"[case] device vs host control, fp64 {kind} batch of {len(models)}", flush=True)
    rd = batch_solve(models, True, monkeypatch, stoptol=1e-8, max_iter=200000)
    rh = batch_solve(models, False, monkeypatch, stoptol=1e-8, max_iter=200000)
    assert all(r.status == "OPTIMAL" for r in rh)
    assert_same(rd, rh, kind)


@pytest.mark.parametrize("kind,max_iter", [("lp", 537), ("lp", 1060), ("csr", 1060), ("diag", 1060)])
def test_fp64_device_control_max_iter(kind, max_iter, monkeypatch, gpu_lib):
    """A max_iter that is neither a check nor a print step: its residuals read
    the bar vectors of the last full update (the update before it is partial,
    algorithm.jl:2895-2900).  1060 follows the check at 1050 with no print step
    between, so an instance restarted at 1050 reads that body's bar vectors and
    the others those of body 1049."""
    models = MODELS[kind]()
    print(f"[case] device vs host control, {kind} at max_iter {max_iter}", flush=True)
    rd = batch_solve(models, True, monkeypatch, stoptol=1e-15, max_iter=max_iter)
    rh = batch_solve(models, False, monkeypatch, stoptol=1e-15, max_iter=max_iter)
    assert sum(r.status == "MAX_ITER" and r.iter == max_iter for r in rh) >= 2
    assert_same(rd, rh, f"max_iter {max_iter}")


def test_fp64_device_control_time_limit(monkeypatch, gpu_lib):
    """time_limit = 0 ends every instance at its first event."""
    models = lp_models()[:4]
    print("[case] device vs host control at time_limit 0", flush=True)
    rd = batch_solve(models, True, monkeypatch, stoptol=1e-14, time_limit=0.0)
    rh = batch_solve(models, False, monkeypatch, stoptol=1e-14, time_limit=0.0)
    assert all(r.status == "TIME_LIMIT" for r in rh)
    assert [r.iter for r in rd] == [r.iter for r in rh]
    assert [r.status for r in rd] == [r.status for r in rh]


def test_fp64_device_control_warm_start(monkeypatch, gpu_lib):
    """Half the instances warm-started: retirements at many different checks."""
    models = lp_models()
    cold = batch_solve(models, False, monkeypatch, stoptol=1e-6)
    rng = np.random.default_rng(5)
    xs = [r.x * (1 + 0.01 * rng.standard_normal(r.x.shape)) if i % 2 == 0 else None for i, r in enumerate(cold)]
    ys = [r.y * (1 + 0.01 * rng.standard_normal(r.y.shape)) if i % 2 == 0 else None for i, r in enumerate(cold)]
    print("[case] device vs host control, warm-started lp batch", flush=True)
    rd = batch_solve(models, True, monkeypatch, warm=(xs, ys), stoptol=1e-6)
    rh = batch_solve(models, False, monkeypatch, warm=(xs, ys), stoptol=1e-6)
    assert_same(rd, rh, "warm")


@pytest.mark.parametrize("kind", ["lp", "portfolio", "regression"])
def test_delta_device_control_agrees_to_rounding(kind, monkeypatch, gpu_lib):
    models = MODELS[kind]()
    print(f"[case] device vs host control, fp32_delta {kind} batch of {len(models)}", flush=True)
    rd = batch_solve(models, True, monkeypatch, stoptol=1e-8, precision="fp32_delta")
    rh = batch_solve(models, False, monkeypatch, stoptol=1e-8, precision="fp32_delta")
    for i, (d, h) in enumerate(zip(rd, rh)):
        print(f"[result] delta {kind}[{i}] device iter {d.iter} (fp32 {d.iter_fp32}) host iter {h.iter} "
              f"(fp32 {h.iter_fp32})", flush=True)
        assert d.status == h.status == "OPTIMAL"
        assert d.residuals < 1e-8
        assert abs(d.iter - h.iter) <= 300
        assert 0 < d.iter_fp32 < d.iter
        assert abs(d.primal_obj - h.primal_obj) <= 1e-7 * (1 + abs(h.primal_obj))



@pytest.mark.parametrize("kind,check_iter,print_frequency,max_iter", [
    ("lp", 30, 7, 480), ("diag", 30, 7, 481), ("lp", 1, -1, 120), ("csr", 45, 10, 450)])
def test_fp64_device_control_event_schedules(kind, check_iter, print_frequency, max_iter, monkeypatch, gpu_lib):
    """Other check and print schedules (every body a check; print steps between
    checks; max_iter on a check, on a print step, and on neither): the device
    loop's events and full updates follow the host loop's exactly."""
    models = MODELS[kind]()
    kw = dict(stoptol=1e-12, check_iter=check_iter, print_frequency=print_frequency, max_iter=max_iter)
    print(f"[case] device vs host control, {kind} {kw}", flush=True)
    rd = batch_solve(models, True, monkeypatch, **kw)
    rh = batch_solve(models, False, monkeypatch, **kw)
    assert_same(rd, rh, f"schedule {check_iter}/{print_frequency}/{max_iter}")
