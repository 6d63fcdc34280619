"""Pins the oracle functions that the reference's own tests leave open — CPU only.

* SOC block-scaling rules (scaling.jl:253-477): every strategy x phase x
  location x size threshold against a hand-written branch table
  (tests/golden/soc_block_rules.json), with block statistics that make every
  outcome distinct (max 16, rms 1: geom 4, maxb 8, taper 2);
* the scaling pipeline (scaling.jl:785-1289): Ruiz, pre-Pock, Pock-Chambolle
  and b/c passes restated densely in numpy with the branch table, against the
  oracle's cumulative row/column factors and scaled data, for all 20 rules;
* check_restart (algorithm.jl:1016-1099): scripted gap sequences, every flag
  and threshold;
* update_sigma! (algorithm.jl:729-990): known-answer cases for the sqrt(b/a)
  proposal, the log blend, every branch of the residual correction and the
  [1e-12, 1e12] clamp; QP proposals against a dense scan of the model;
* compute_M_norm! (algorithm.jl:1302-1359): the value recomputed in numpy from
  dx, dy, dw, A, Q, sigma and lambda for noQ, diagonal, CSR and factored Q.
"""
import ctypes as C
import json
import math
import os

import numpy as np
import pytest
import scipy.sparse as sp

import hprsocp
from hprsocp import _abi, synthetic
from kat_models import params

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "soc_block_rules.json")
TABLE = json.load(open(GOLD))
PHASES = TABLE["phases"]
LOCS = TABLE["locations"]


def outcome_value(name, bm, rms):
    return {"max": bm, "rms": rms, "geom": math.sqrt(bm * rms),
            "maxb": (bm ** 3 * rms) ** 0.25, "taper": (bm * rms ** 3) ** 0.25}[name]


def table_outcome(strategy, phase, location, L, cone_count):
    """Outcome name of the hand-written table for one case."""
    if L <= 0:
        return None
    rule = TABLE["rules"][strategy]
    by_loc = rule.get("any") or rule.get(location) or rule["other"]
    o = by_loc.get(phase, by_loc["rest"])
    if isinstance(o, dict):
        many = cone_count >= TABLE["thresholds"]["many_cones"] and L <= TABLE["thresholds"]["adaptive"]
        o = o["many_small"] if many else o["else"]
    if isinstance(o, list):
        for bound, name in o:
            if bound is None or L <= bound:
                return name
    return o


def table_value(strategy, phase, location, L, bm, rms, cone_count=1):
    o = table_outcome(strategy, phase, location, L, cone_count)
    return 1.0 if o is None else outcome_value(o, bm, rms)


# ------------------------------------------------------------- block rules
def _oracle_rule(lib):
    fn = lib.oracle_soc_block_scale_value
    fn.restype = C.c_double
    fn.argtypes = [C.c_int32, C.c_double, C.c_double, C.c_int32, C.c_int32, C.c_int32, C.c_int32]
    return fn


@pytest.mark.parametrize("strategy", sorted(_abi.SOC_STRATEGIES))
def test_block_rule_table(oracle_lib, strategy):
    fn = _oracle_rule(oracle_lib)
    code = _abi.SOC_STRATEGIES[strategy]
    seen = set()
    for bm, rms in ((16.0, 1.0), (2.25, 1.21)):
        for ph, phc in PHASES.items():
            for loc, locc in LOCS.items():
                for L in (0, 1, 8, 9, 512, 513, 4000):
                    for cc in (1, 99999, 100000):
                        want = table_value(strategy, ph, loc, L, bm, rms, cc)
                        got = fn(L, bm, rms, code, phc, locc, cc)
                        assert got == pytest.approx(want, rel=1e-15), (strategy, ph, loc, L, cc, bm, got, want)
                        seen.add(table_outcome(strategy, ph, loc, L, cc))
    # the grid reaches every branch the table names for this strategy
    names = set()

    def walk(o):
        if isinstance(o, str):
            names.add(o)
        elif isinstance(o, list):
            for _, x in o:
                names.add(x)
        elif isinstance(o, dict):
            for k, v in o.items():
                if k != "line":
                    walk(v)
    walk(TABLE["rules"][strategy])
    assert names <= seen, names - seen


def test_block_rule_distinct_outcomes():
    """With max 16 and rms 1 every outcome is a different number, so a swapped
    branch cannot pass by coincidence."""
    vals = [outcome_value(o, 16.0, 1.0) for o in ("max", "rms", "geom", "maxb", "taper")]
    assert len(set(vals)) == 5 and vals == [16.0, 1.0, 4.0, 8.0, 2.0]


def test_default_strategy_rows_in_table():
    """The default :phase_taper (structs.jl default; GPU kernel kernels.jl:6869-6906)."""
    t = "phase_taper"
    assert table_value(t, "ruiz", "constraint", 8, 16, 1) == 16
    assert table_value(t, "ruiz", "variable", 9, 16, 1) == 1
    assert table_value(t, "pock", "constraint", 8, 16, 1) == 1
    assert table_value(t, "pock", "constraint", 9, 16, 1) == 4
    assert table_value(t, "pre_pock", "variable", 3, 16, 1) == 4
    assert table_value(t, "generic", "generic", 3, 16, 1) == 4


# ------------------------------------------------------------- scaling pipeline
def numpy_scaling(model, strategy, ruiz_iterations=3, use_ruiz=True, use_pc=True, use_bc=True,
                  bc_norm="l2"):
    """Dense restatement of scaling! (scaling.jl:785-1289, CPU branches) with
    the branch table for the cone blocks.  Returns the cumulative factors and
    the scaled data."""
    A = model.A.toarray().astype(float)
    m, n = A.shape
    Q = np.zeros((n, n)) if model.Q is None else sp.csr_matrix(model.Q).toarray().astype(float)
    c = np.array(model.c, float)
    AL, AU = np.array(model.AL, float), np.array(model.AU, float)
    l, u = np.array(model.l, float), np.array(model.u, float)
    m1 = model.number_linear_con
    rhs_full = np.zeros(m)
    rhs_full[m1:] = np.asarray(model.soc_rhs, float)
    ci = np.asarray(model.SOC_con_idx) - 1
    vi = np.asarray(model.SOC_var_idx) - 1
    row, col = np.ones(m), np.ones(n)

    def blocks(t, idx, phase, loc):
        k = len(idx) - 1
        for j in range(k):
            s, e = idx[j], idx[j + 1]
            if e > s:
                blk = t[s:e]
                t[s:e] = table_value(strategy, phase, loc, e - s, blk.max(), math.sqrt(np.mean(blk ** 2)), k)
        return t

    def one_pass(kind, phase):
        nonlocal A, Q, c, AL, AU, l, u, rhs_full, row, col
        if kind == "max":
            tc = np.sqrt(np.maximum(np.abs(A).max(axis=0, initial=0.0), np.abs(Q).max(axis=0, initial=0.0)))
            tr = np.sqrt(np.abs(A).max(axis=1, initial=0.0))
        else:
            tc = np.sqrt(np.abs(A).sum(axis=0) + np.abs(Q).sum(axis=0))
            tr = np.sqrt(np.abs(A).sum(axis=1))
        tc[tc == 0] = 1.0
        tr[tr == 0] = 1.0
        tr = np.clip(blocks(tr, ci, phase, "constraint"), 1e-4, 1e4)
        tc = np.clip(blocks(tc, vi, phase, "variable"), 1e-4, 1e4)
        row, col = row * tr, col * tc
        Q = Q / tc[:, None] / tc[None, :]
        A = A / tr[:, None] / tc[None, :]
        c = c / tc
        AL, AU, rhs_full = AL / tr, AU / tr, rhs_full / tr
        l, u = l * tc, u * tc

    if use_ruiz:
        for _ in range(ruiz_iterations):
            one_pass("max", "ruiz")
    if strategy in TABLE["pre_pock_strategies"]["names"]:
        for _ in range(2):
            one_pass("max", "pre_pock")
    if use_pc:
        one_pass("sum", "pock")
    b_scale = c_scale = 1.0
    if use_bc:
        def fin(v):
            v = v[np.isfinite(v)]
            if bc_norm == "linf":
                return np.abs(v).max(initial=0.0)
            if v.size == 0:
                return 0.0
            return math.sqrt(np.sum(v * v) / (v.size if bc_norm == "rms" else 1))

        def vec(v):
            if v.size == 0:
                return 0.0
            if bc_norm == "linf":
                return np.abs(v).max()
            return math.sqrt(np.sum(v * v) / (v.size if bc_norm == "rms" else 1))
        soc = vec(rhs_full[m1:]) if len(ci) > 1 else 0.0
        b_scale = float(np.clip(max(1.0, fin(AL), fin(AU), soc), 1e-4, 1e4))
        hat_c = float(np.clip(max(1.0, vec(c)), 1e-4, 1e4))
        c_scale = float(np.clip(hat_c, b_scale / 1e2, b_scale / 1e-2))
        Q = Q * (b_scale / c_scale)
        AL, AU, rhs_full = AL / b_scale, AU / b_scale, rhs_full / b_scale
        c = c / c_scale
        l, u = l / b_scale, u / b_scale
    return dict(row_norm=row, col_norm=col, A=A, Q=Q, c=c, AL=AL, AU=AU, l=l, u=u,
                soc_rhs_full=rhs_full, b_scale=b_scale, c_scale=c_scale)


def _scaling_model():
    # cone blocks on both sides straddle the size thresholds (8/9 rows); rows and
    # columns of very different magnitude so that max, rms and geom differ
    m = synthetic.make_small_qp(seed=31, q="diag", n1=14, var_sizes=(3, 8, 9, 20), m1=9,
                                con_sizes=(2, 8, 9, 25), density=0.45)
    A = m.A.tocsr().astype(float)
    rng = np.random.default_rng(4)
    D = sp.diags(np.exp(rng.uniform(-2.5, 2.5, A.shape[0])))
    E = sp.diags(np.exp(rng.uniform(-2.5, 2.5, A.shape[1])))
    m.A = sp.csr_matrix(D @ A @ E)
    return m


@pytest.mark.parametrize("strategy", sorted(_abi.SOC_STRATEGIES))
def test_scaling_pipeline_numpy(oracle_lib, strategy):
    model = _scaling_model()
    want = numpy_scaling(model, strategy, ruiz_iterations=3)
    p = params(soc_block_scaling_strategy=strategy, ruiz_iterations=3)
    with hprsocp.Solver(model, p, lib=oracle_lib, prefix="oracle_") as o:
        for k in ("row_norm", "col_norm", "c", "AL", "AU", "l", "u", "soc_rhs_full"):
            np.testing.assert_allclose(o.get(k), want[k], rtol=1e-12, atol=0, err_msg=k)
        A = model.A.tocsr()
        As = sp.csr_matrix((o.scaled_A(), A.indices, A.indptr), shape=A.shape).toarray()
        np.testing.assert_allclose(As, want["A"], rtol=1e-12, atol=1e-300)
        np.testing.assert_allclose(o.get("diag_Q"), np.diag(want["Q"]), rtol=1e-12)
        s = o.scalars()
        assert s["b_scale"] == pytest.approx(want["b_scale"], rel=1e-13)
        assert s["c_scale"] == pytest.approx(want["c_scale"], rel=1e-13)
    # the factors really depend on the rule: the cone blocks are not all max
    if strategy != "max":
        ref = numpy_scaling(model, "max", ruiz_iterations=3)
        assert np.max(np.abs(want["row_norm"] / ref["row_norm"] - 1)) > 1e-3


@pytest.mark.parametrize("bc_norm", ["l2", "rms", "linf"])
@pytest.mark.parametrize("stages", [(True, False), (False, True), (False, False)])
def test_scaling_stages_numpy(oracle_lib, bc_norm, stages):
    ruiz, pc = stages
    model = _scaling_model()
    want = numpy_scaling(model, "phase_taper", ruiz_iterations=10, use_ruiz=ruiz, use_pc=pc,
                         bc_norm=bc_norm)
    p = params(use_Ruiz_scaling=ruiz, use_Pock_Chambolle_scaling=pc, bc_scaling_norm_type=bc_norm)
    with hprsocp.Solver(model, p, lib=oracle_lib, prefix="oracle_") as o:
        for k in ("row_norm", "col_norm", "c", "AL", "AU", "l", "u"):
            np.testing.assert_allclose(o.get(k), want[k], rtol=1e-12, err_msg=k)
        s = o.scalars()
        assert s["b_scale"] == pytest.approx(want["b_scale"], rel=1e-13)
        assert s["c_scale"] == pytest.approx(want["c_scale"], rel=1e-13)


# ------------------------------------------------------------- check_restart
def _restart(lib, **kw):
    r = _abi.RestartInfo()
    lib.oracle_init_restart(C.byref(r))
    for k, v in kw.items():
        setattr(r, k, v)
    return r


def _check(lib, r, it, sigma=1.0, kkt=1.0, check_iter=150):
    lib.oracle_check_restart(C.byref(r), it, check_iter, sigma, kkt)
    return r


def test_initialize_restart(oracle_lib):
    r = _restart(oracle_lib)
    assert r.first_restart == 1 and r.restart_flag == 0
    assert math.isinf(r.last_gap) and math.isinf(r.current_gap) and math.isinf(r.save_gap)
    assert math.isinf(r.best_gap) and math.isinf(r.best_kkt) and math.isinf(r.weighted_norm)
    assert r.best_sigma == 1.0 and r.inner == 0 and r.times == 0 and r.best_iter == 0
    assert r.sigma_correction_active == 0 and r.sigma_correction_hold == 0 and r.sigma_correction_dir == 0


def test_check_restart_first(oracle_lib):
    L = oracle_lib
    r = _check(L, _restart(L, current_gap=0.7), 0)
    assert r.restart_flag == 0 and r.first_restart == 1
    r = _check(L, _restart(L, current_gap=0.7), 149)
    assert r.restart_flag == 0 and r.first_restart == 1
    r = _check(L, _restart(L, current_gap=0.7), 150, sigma=3.0, kkt=0.01)
    assert r.restart_flag == 1 and r.first_restart == 0
    assert (r.weighted_norm, r.best_gap, r.best_kkt, r.best_sigma) == (0.7, 0.7, 0.01, 3.0)
    assert math.isinf(r.save_gap) and r.sufficient == 0
    # after the first restart, iterations off the check grid change nothing
    before = r.as_dict()
    r = _check(L, r, 225)
    after = r.as_dict()
    before["restart_flag"] = 0
    assert after == before


def _after_first(L, **kw):
    r = _restart(L, first_restart=0, best_gap=10.0, best_kkt=10.0, best_sigma=5.0)
    for k, v in kw.items():
        setattr(r, k, v)
    return r


@pytest.mark.parametrize("cur,flag", [(0.36, 1), (0.36000001, 0), (0.1, 1)])
def test_check_restart_sufficient(oracle_lib, cur, flag):
    r = _check(oracle_lib, _after_first(oracle_lib, last_gap=1.0, current_gap=cur, inner=10), 300)
    assert r.restart_flag == flag and r.sufficient == flag
    assert r.save_gap == cur


def test_check_restart_necessary(oracle_lib):
    L = oracle_lib
    r = _check(L, _after_first(L, last_gap=1.0, save_gap=0.5, current_gap=0.6, inner=10), 300)
    assert r.restart_flag == 2 and r.necessary == 1 and r.sufficient == 0
    r = _check(L, _after_first(L, last_gap=1.0, save_gap=0.6, current_gap=0.6, inner=10), 300)
    assert r.restart_flag == 0  # current_gap must exceed save_gap
    r = _check(L, _after_first(L, last_gap=1.0, save_gap=0.5, current_gap=0.81, inner=10), 300)
    assert r.restart_flag == 0  # 0.81 > 0.8 last_gap
    # sufficient and necessary together: the later test wins, both counters move
    r = _check(L, _after_first(L, last_gap=1.0, save_gap=0.2, current_gap=0.3, inner=10), 300)
    assert r.restart_flag == 2 and r.sufficient == 1 and r.necessary == 1


@pytest.mark.parametrize("inner,it,flag", [(500, 3000, 3), (499, 3000, 0), (500, 3450, 0), (518, 3450, 3)])
def test_check_restart_long(oracle_lib, inner, it, flag):
    # long: inner >= max(500, 0.15 iter) and current_gap > 0.9 save_gap
    r = _check(oracle_lib, _after_first(oracle_lib, last_gap=1.0, save_gap=1.0, current_gap=1.0, inner=inner),
               it)
    assert r.restart_flag == flag and r.long_restarts == (1 if flag else 0)


def test_check_restart_long_10000(oracle_lib):
    L = oracle_lib
    r = _check(L, _after_first(L, last_gap=1.0, save_gap=0.8, current_gap=0.75, inner=10000), 150000)
    assert r.restart_flag == 3
    r = _check(L, _after_first(L, last_gap=1.0, save_gap=0.8, current_gap=0.76, inner=10000), 150000)
    assert r.restart_flag == 0  # 0.76 > 0.75 last_gap
    r = _check(L, _after_first(L, last_gap=1.0, save_gap=0.9, current_gap=0.75, inner=10000), 150000)
    assert r.restart_flag == 0  # 0.75 <= 0.9 save_gap * 0.9 = 0.81


def test_check_restart_negative_gap(oracle_lib):
    r = _check(oracle_lib, _after_first(oracle_lib, last_gap=1.0, current_gap=-3.0, inner=10), 300)
    assert r.current_gap == 1e-6 and r.restart_flag == 1 and r.save_gap == 1e-6


def test_check_restart_best_refresh(oracle_lib):
    L = oracle_lib
    base = dict(last_gap=1.0, current_gap=0.9, inner=10, best_gap=0.5, best_kkt=1.0, best_sigma=5.0, best_iter=3)
    r = _check(L, _after_first(L, **base), 300, sigma=2.0, kkt=0.99)   # improved KKT
    assert (r.best_gap, r.best_kkt, r.best_sigma, r.best_iter) == (0.9, 0.99, 2.0, 0)
    r = _check(L, _after_first(L, **dict(base, current_gap=0.4)), 300, sigma=2.0, kkt=1.01)  # compatible gap
    assert (r.best_gap, r.best_kkt, r.best_sigma, r.best_iter) == (0.4, 1.01, 2.0, 0)
    r = _check(L, _after_first(L, **dict(base, current_gap=0.4)), 300, sigma=2.0, kkt=1.0101)
    assert r.best_sigma == 5.0 and r.best_iter == 3  # 1.0101 > 1.01 best_kkt
    r = _check(L, _after_first(L, **dict(base, best_iter=10)), 300, sigma=2.0, kkt=1.2)  # periodic
    assert (r.best_sigma, r.best_iter, r.best_kkt) == (2.0, 0, 1.2)
    r = _check(L, _after_first(L, **dict(base, best_iter=9)), 300, sigma=2.0, kkt=1.2)
    assert r.best_sigma == 5.0 and r.best_iter == 9


# ------------------------------------------------------------- update_sigma!
def _tiny_noq():
    return synthetic.make_small_qp(seed=41, q="none", n1=6, var_sizes=(3,), m1=2, con_sizes=(3,), density=0.6)


def _set_diffs(o, dx, dy, dw=None, daty=None):
    n, m = o.n, o.m
    o.set("last_x", np.zeros(n))
    o.set("x_bar", dx)
    o.set("last_y", np.zeros(m))
    o.set("y_bar", dy)
    if dw is not None:
        o.set("last_w", np.zeros(n))
        o.set("w_bar", dw)
    if daty is not None:
        o.set("last_ATy", np.zeros(n))
        o.set("ATy_bar", daty)


def _unit(k, norm2):
    v = np.zeros(k)
    v[0] = math.sqrt(norm2)
    return v


# (restart fields, residuals, sigma_old) -> expected sigma, correction state.
# Proposal: lambda_A = 2, |dy|^2 = 0.5, |dx|^2 = 4 -> a = 1, b = 4, sigma_model = 2.
SIGMA_CASES = {
    "reference_off": (dict(best_kkt=1e-3, best_gap=1.0, current_gap=1.0, best_sigma=100.0),
                      dict(kkt=1.0, err_Rp=1e-3, err_Rd=1e-3), 1.0, 2.0, (0, -1, 0)),
    "blend_theta_1": (dict(best_kkt=1e-3, best_gap=1.0, current_gap=1.0, best_sigma=100.0),
                      dict(kkt=1e-3, err_Rp=1e-3, err_Rd=1e-3), 1.0, 2.0, (0, -1, 0)),
    "blend_theta_clamped": (dict(best_kkt=1e-3, best_gap=1.0, current_gap=1.0, best_sigma=100.0),
                            dict(kkt=0.25e-3, err_Rp=1e-3, err_Rd=1e-3), 1.0, 2.0, (0, -1, 0)),
    "blend_pr_4": (dict(best_kkt=1e-3, best_gap=1.0, current_gap=4.0, best_sigma=100.0),
                   dict(kkt=4e-3, err_Rp=1e-3, err_Rd=1e-3), 1.0,
                   2.0 ** math.exp(-0.9) * 100.0 ** (1 - math.exp(-0.9)), (0, -1, 0)),
    "extreme_dual_worse": (dict(best_kkt=1e-3, best_gap=1.0, current_gap=1.0),
                           dict(kkt=1.0, err_Rp=1e-10, err_Rd=1e-3), 1.0, 200.0, (1, 0, 1)),
    "extreme_primal_worse": (dict(best_kkt=1e-3, best_gap=1.0, current_gap=1.0),
                             dict(kkt=1.0, err_Rp=1e-3, err_Rd=1e-10), 1.0, 0.02, (1, 0, -1)),
    "extreme_push_floor": (dict(best_kkt=1e-3, best_gap=1.0, current_gap=1.0),
                           dict(kkt=1.0, err_Rp=1e-3, err_Rd=1e-10), 0.01, 0.01 / 1.5, (1, 0, -1)),
    "one_side_finished": (dict(best_kkt=1e-3, best_gap=1.0, current_gap=1.0),
                          dict(kkt=1.0, err_Rp=3e-8, err_Rd=4e-9), 1.0, 1.0 / 1.5, (1, 0, -1)),
    "one_side_no_push": (dict(best_kkt=1e-3, best_gap=1.0, current_gap=1.0),
                         dict(kkt=1.0, err_Rp=3e-8, err_Rd=4e-9), 4.0, 2.0 * math.sqrt(4.0 / 30.0), (1, 0, -1)),
    "keep_active": (dict(best_kkt=1e-3, best_gap=1.0, current_gap=1.0, sigma_correction_active=1,
                         sigma_correction_dir=-1),
                    dict(kkt=1.0, err_Rp=5e-8, err_Rd=7e-9), 2.0, 2.0 * math.sqrt(0.14), (1, 0, -1)),
    "keep_floor_ratio": (dict(best_kkt=1e-3, best_gap=1.0, current_gap=1.0, sigma_correction_active=1,
                              sigma_correction_dir=1),
                         dict(kkt=1.0, err_Rp=7e-9, err_Rd=1.2e-8), 0.1, 2.0 * math.sqrt(2.0), (1, 0, 1)),
    "keep_released": (dict(best_kkt=1e-3, best_gap=1.0, current_gap=1.0, sigma_correction_active=1,
                           sigma_correction_dir=-1),
                      dict(kkt=1.0, err_Rp=5e-8, err_Rd=1e-8), 2.0, 2.0, (0, -1, 0)),
    "keep_edge_in": (dict(best_kkt=1e-3, best_gap=1.0, current_gap=1.0, sigma_correction_active=1,
                          sigma_correction_dir=-1),
                     dict(kkt=1.0, err_Rp=5e-8, err_Rd=7.9e-9), 2.0, 2.0 * math.sqrt(7.9e-9 / 5e-8), (1, 0, -1)),
    "keep_edge_out": (dict(best_kkt=1e-3, best_gap=1.0, current_gap=1.0, sigma_correction_active=1,
                           sigma_correction_dir=-1),
                      dict(kkt=1.0, err_Rp=5e-8, err_Rd=8.1e-9), 2.0, 2.0, (0, -1, 0)),
    "keep_by_ratio": (dict(best_kkt=1e-3, best_gap=1.0, current_gap=1.0, sigma_correction_active=1,
                           sigma_correction_dir=-1),
                      dict(kkt=1.0, err_Rp=1e-2, err_Rd=5e-8), 1.0, 2.0 * 0.1, (1, 0, -1)),  # log ratio clamped to -log 100
    "ratio_below_keep": (dict(best_kkt=1e-3, best_gap=1.0, current_gap=1.0, sigma_correction_active=1,
                              sigma_correction_dir=-1),
                         dict(kkt=1.0, err_Rp=4e-3, err_Rd=5e-8), 1.0, 2.0, (0, -1, 0)),
    "both_finished": (dict(best_kkt=1e-3, best_gap=1.0, current_gap=1.0, sigma_correction_active=1,
                           sigma_correction_dir=1),
                      dict(kkt=1.0, err_Rp=5e-9, err_Rd=1e-12), 1.0, 2.0, (0, -1, 0)),
}


@pytest.mark.parametrize("case", sorted(SIGMA_CASES))
def test_update_sigma_kat(oracle_lib, case):
    fields, res, sigma_old, want, corr = SIGMA_CASES[case]
    model = _tiny_noq()
    with hprsocp.Solver(model, params(), lib=oracle_lib, prefix="oracle_") as o:
        o.set_scalars(sigma=sigma_old, lambda_A=2.0, lambda_Q=0.0)
        _set_diffs(o, _unit(o.n, 4.0), _unit(o.m, 0.5))
        r = _restart(oracle_lib, first_restart=0, restart_flag=1, best_iter=4, **fields)
        o.update_sigma(r, res)
        assert o.scalars()["sigma"] == pytest.approx(want, rel=1e-14), case
        assert (r.sigma_correction_active, r.sigma_correction_hold, r.sigma_correction_dir) == corr
        assert r.best_iter == 5


@pytest.mark.parametrize("which,want", [("low", 1e-12), ("high", 1e12)])
def test_update_sigma_clamp(oracle_lib, which, want):
    model = _tiny_noq()
    with hprsocp.Solver(model, params(), lib=oracle_lib, prefix="oracle_") as o:
        o.set_scalars(sigma=1.0, lambda_A=1.0, lambda_Q=0.0)
        if which == "low":   # b = |dx|^2 = 0 -> EPS; a = 1e10 -> sqrt(1e-26) = 1e-13
            _set_diffs(o, np.zeros(o.n), _unit(o.m, 1e10))
        else:                # a = 0 -> EPS; b = 1e10 -> 1e13
            _set_diffs(o, _unit(o.n, 1e10), np.zeros(o.m))
        r = _restart(oracle_lib, first_restart=0, restart_flag=2, best_kkt=1e-3, best_gap=1.0, current_gap=1.0)
        o.update_sigma(r, dict(kkt=1.0, err_Rp=1e-3, err_Rd=1e-3))
        assert o.scalars()["sigma"] == want


def test_update_sigma_no_flag(oracle_lib):
    with hprsocp.Solver(_tiny_noq(), params(), lib=oracle_lib, prefix="oracle_") as o:
        o.set_scalars(sigma=0.7, lambda_A=2.0)
        _set_diffs(o, _unit(o.n, 4.0), _unit(o.m, 0.5))
        r = _restart(oracle_lib, first_restart=0, restart_flag=0, best_iter=4)
        before = r.as_dict()
        o.update_sigma(r, dict(kkt=1.0, err_Rp=1e-3, err_Rd=1e-3))
        assert o.scalars()["sigma"] == 0.7 and r.as_dict() == before


def _qp_unscaled(q, seed):
    m = synthetic.make_small_qp(seed=seed, q=q, n1=8, var_sizes=(3, 4), m1=3, con_sizes=(3,), density=0.5)
    return m


NOSCALE = dict(use_Ruiz_scaling=False, use_Pock_Chambolle_scaling=False, use_bc_scaling=False)


def _qmat(model):
    """Q as the oracle applies it (F'F + D for the factored form)."""
    if isinstance(model.Q, tuple):
        _, F, D = model.Q
        F = sp.csr_matrix(F)
        D = np.zeros(model.A.shape[1]) if D is None else np.asarray(D)
        return (F.T @ F).toarray() + np.diag(D)
    return sp.csr_matrix(model.Q).toarray()


@pytest.mark.parametrize("q", ["diag", "csr"])
def test_update_sigma_qp_model_minimiser(oracle_lib, q):
    """QP proposal: sigma_model minimises the surrogate of algorithm.jl:781-829 (dots
    recomputed in numpy); with the reference blend off and balanced residuals,
    sigma_new = sigma_model.  Checked against a dense log-scan of the surrogate."""
    model = _qp_unscaled(q, 51 if q == "diag" else 52)
    rng = np.random.default_rng(9)
    with hprsocp.Solver(model, params(**NOSCALE), lib=oracle_lib, prefix="oracle_") as o:
        n, mm = o.n, o.m
        lam_A, lam_Q = 3.0, 2.5
        o.set_scalars(sigma=1.0, lambda_A=lam_A, lambda_Q=lam_Q)
        dx, dy, dw, daty = rng.normal(size=n), rng.normal(size=mm), rng.normal(size=n), rng.normal(size=n)
        _set_diffs(o, dx, dy, dw, daty)
        Qm = _qmat(model)
        dqw, qatdy = Qm @ dw, Qm @ daty
        b = dx @ dx
        a = lam_A * dy @ dy - 2 * dqw @ daty
        diag = np.diag(Qm)
        if q == "diag":
            a += dqw @ dqw
            f = lambda x: a * x + b / x + x * x * np.sum(daty * qatdy / (1 + x * diag))  # noqa: E731
        else:
            a += lam_Q * dw @ dqw
            cc = daty @ qatdy
            f = lambda x: a * x + b / x + cc * x * x / (1 + lam_Q * x)  # noqa: E731
        a = max(a, 1e-16)
        xs = np.exp(np.linspace(math.log(1e-6), math.log(1e6), 20001))
        fx = np.array([f(x) for x in xs])
        k = int(np.argmin(fx))
        r = _restart(oracle_lib, first_restart=0, restart_flag=1, best_kkt=1e-3, best_gap=1.0, current_gap=1.0)
        o.update_sigma(r, dict(kkt=1.0, err_Rp=1e-3, err_Rd=1e-3))
        s = o.scalars()["sigma"]
        assert xs[max(k - 1, 0)] <= s <= xs[min(k + 1, len(xs) - 1)], (s, xs[k])
        assert f(s) <= fx[k] + 1e-12 * abs(fx[k])


# ------------------------------------------------------------- compute_M_norm!
def mnorm_numpy(A, Qm, dx, dy, dw, sigma, lam_A, lam_Q, qdiag):
    """compute_M_norm! (algorithm.jl:1302-1359) in numpy."""
    M1, M2, M3 = 0.0, dx @ dx / sigma, 0.0
    has_q = Qm is not None and np.any(Qm != 0)
    m = A.shape[0]
    if has_q:
        dqw = Qm @ dw
        M2 -= 2 * dqw @ dx
    if m > 0:
        M1 = sigma * lam_A * dy @ dy
        atdy = A.T @ dy
        if has_q:
            qatdy = Qm @ atdy
            M1 -= 2 * sigma * dqw @ atdy
        M2 += 2 * atdy @ dx
        if has_q:
            if qdiag:
                fact_M = sigma * sigma / (1 + sigma * np.diag(Qm))
                M3 = (atdy * fact_M) @ qatdy
                M1 += sigma * dqw @ dqw
            else:
                M3 = sigma ** 2 / (1 + sigma * lam_Q) * atdy @ qatdy
                M1 += sigma * lam_Q * dw @ dqw
    elif has_q and not qdiag:
        M1 = sigma * lam_Q * dw @ dqw
    M2 += max(M1, 0.0)
    return math.sqrt(max(M2, 0.0) + max(M3, 0.0))


@pytest.mark.parametrize("q", ["none", "diag", "csr", "factored"])
def test_mnorm_numpy(oracle_lib, q):
    model = _qp_unscaled(q, {"none": 61, "diag": 62, "csr": 63, "factored": 64}[q])
    rng = np.random.default_rng(17)
    with hprsocp.Solver(model, params(**NOSCALE), lib=oracle_lib, prefix="oracle_") as o:
        n, mm = o.n, o.m
        sigma, lam_A, lam_Q = 0.8, 3.0, 2.5
        o.set_scalars(sigma=sigma, lambda_A=lam_A, lambda_Q=lam_Q)
        dx, dy, dw = rng.normal(size=n), rng.normal(size=mm), rng.normal(size=n)
        o.set("dx", dx)
        o.set("dy", dy)
        o.set("dw", dw)
        Qm = None if q == "none" else _qmat(model)
        qdiag = bool(o.scalars()["Q_is_diag"])
        if q != "none":  # an empty Q counts as diagonal (check_Q_diagonal)
            assert qdiag == (q == "diag")
        want = mnorm_numpy(model.A.toarray(), Qm, dx, dy, dw, sigma, lam_A, lam_Q, qdiag)
        assert o.mnorm() == pytest.approx(want, rel=1e-12)
        # every term matters: dropping the A^T dy cross term changes the value
        alt = mnorm_numpy(np.zeros_like(model.A.toarray()), Qm, dx, dy, dw, sigma, lam_A, lam_Q, qdiag)
        assert abs(alt - want) > 1e-6 * want


def test_mnorm_numpy_no_rows(oracle_lib):
    """m = 0 with a general Q: only sigma lambda_Q <dw, Q dw> enters M1."""
    rng = np.random.default_rng(3)
    n = 5
    F = sp.random(3, n, density=0.8, random_state=2, data_rvs=rng.standard_normal)
    Qm = (F.T @ F).toarray() + np.eye(n) * 0.1
    Qm[0, 1] = Qm[1, 0] = Qm[0, 1] + 0.3
    model = hprsocp.build_qp(sp.csr_matrix((0, n)), rng.normal(size=n), [], [], -np.ones(n), np.ones(n), [], 0,
                             n, [0], [n], Q=sp.csr_matrix(Qm))
    with hprsocp.Solver(model, params(**NOSCALE), lib=oracle_lib, prefix="oracle_") as o:
        sigma, lam_Q = 1.3, 4.0
        o.set_scalars(sigma=sigma, lambda_A=1.0, lambda_Q=lam_Q)
        dx, dw = rng.normal(size=n), rng.normal(size=n)
        o.set("dx", dx)
        o.set("dw", dw)
        want = mnorm_numpy(np.zeros((0, n)), Qm, dx, np.zeros(0), dw, sigma, 1.0, lam_Q, False)
        assert o.mnorm() == pytest.approx(want, rel=1e-12)
