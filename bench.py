#!/usr/bin/env python3
"""Benchmark of the HPR iteration on one MI355X.

Single problem (default --config c1): the large QP-SOCP, n = 1e6 boxed
variables, m_A = 5e5 linear rows with log-normal row lengths (mean 20) plus
10,000 constraint cones of sizes U[3,201], Q = F'F + diag(U(0,1)) with F
1e5 x 1e6 at 5 nnz per column (factored operator form); seeded synthetic data.
A "step" is one HPR iteration (main update, partial outputs) over the resident
instance; `value` = FP64 HPR iterations/s.  Also measured: iterations/s of the
two FP32 forms (fp32_delta: FP32 differences from the FP64 restart anchors;
fp32_refine: FP32 state, FP64 checks, FP64 takeover), one solve to KKT 1e-8 in
FP64 and in fp32_delta (the 1e-6 milestone of the same run gives the time to
1e-6), the roofline of every launch group in each precision (HIP events on the
solver's stream), the SpMV gather rates against the probe's FP64 and FP32
gather ceilings, the autotuner's launch plans, and the CPU oracle on a bounded
sample.

Batches (part of the default run unless --no-batch; alone with --config c3/c4):
  C3: 512 portfolio SOCPs, n log-uniform in [200, 20000], in the three
      precisions, against all 512 sequential single-problem FP64 solves;
  C4: 64 robust-regression QP-SOCPs, d ~ U[1e4, 2e5], half warm-started, in
      FP64 and fp32_delta at 1e-6 and 1e-8, against the 64 sequential FP64
      solves at each tolerance.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c1|c2|c0|c3|c4] [--scale s]
"""
import argparse
import ctypes as C
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "hpr-socp_amd")
sys.path.insert(0, PKG)

METRIC = "time-to-1e-6 KKT & HPR iters/s (FP64 vs FP32+refine); batched solves/s"
SINGLE_PRECS = ["fp64", "fp32_delta", "fp32_refine"]
SOLVE_PRECS = ["fp64", "fp32_delta"]  # fp32_refine: iteration rate only (--refine-solve adds its solve)
BATCH_PRECS = ["fp64", "fp32_refine", "fp32_delta"]
C4_PRECS = ["fp64", "fp32_delta"]
T_START = time.time()
BUDGET_S = 450.0  # optional legs are skipped past this wall time (the run must end well within 10 minutes)


def precisions(args, every):
    if args.precision == "all":
        return every
    want = ["fp64", "fp32_delta"] if args.precision == "both" else [args.precision]
    return [p for p in every if p in want] or want


PEAK_HBM_GBS = 8000.0

# HBM bytes per launch of the dominant launch group, from two rocprofv3 --pmc
# passes (FETCH_SIZE, WRITE_SIZE) over this script, each scaled by what the same
# counter reads for the 1 GiB plain copy of hprsocp_probe_copy (read factor 2.0,
# write factor 1.0).  Source: profiles/c1_fp64_iteration_hbm_traffic.txt.
# Keyed by (config, precision, group, SpMV form that the autotuner kept for the
# group's operator (A for dual, A^T for w2), kernel revision); a run whose form
# differs reports null instead of a stale figure, and KERNEL_REV changes
# whenever the launch group's kernels change.
KERNEL_REV = "rev6"
MEASURED_TRAFFIC = {
    # k_spmv_sorted_fused (A, slice-fused) + k_tile<OpDual, slices>
    ("c1", "fp64", "dual", "sorted8-fused", "rev6"): 496.3e6 + 81.0e6,
    # k_spmv_sorted (A, 8 slice partials) + k_tile<OpDual, slices>
    ("c1", "fp64", "dual", "sorted8", "rev6"): 541.6e6 + 172.3e6,
    # the same for A^T + k_tile<OpW2, slices>
    ("c1", "fp64", "w2", "sorted8-fused", "rev6"): 554.0e6 + 64.5e6,
    ("c1", "fp64", "w2", "sorted8", "rev6"): 508.3e6 + 120.5e6,
}
GROUP_OPERATOR = {"dual": "A", "w2": "A^T"}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes_dims(m, n, nnz, m1, n1, Q, s=8, fused_qw=False):
    """Compulsory HBM bytes per partial HPR iteration, per launch group (s = value bytes).

    Each matrix is read once per SpMV (s + 4 bytes per nonzero, 4 bytes per row
    pointer), each vector once per kernel; gathered SpMV inputs count once.
    fused_qw: the primal step runs in the F^T product's kernel (groups "Fw" and
    "FTw_primal"; Q w is written once and not read back, w is read once)."""
    m2 = m - m1
    out = {}
    if isinstance(Q, tuple):
        F = Q[1]
        mf, nnzf = F.shape[0], F.nnz
        f1 = nnzf * (s + 4) + 4 * (mf + 1) + s * (n + mf)             # F v
        f2 = nnzf * (s + 4) + 4 * (n + 1) + s * (mf + 3 * n)          # F'(.) + D v
        if fused_qw:
            out["Fw"] = f1
            out["FTw_primal"] = f2 + s * 9 * n
        else:
            out["Qw"] = f1 + f2
            out["qp_primal"] = s * (8 * n + 3 * n)
        out["tempv"] = f1 + f2 + s * 3 * n                            # + x_hat, qw, tempv
        out["dual"] = nnz * (s + 4) + 4 * (m + 1) + s * n + s * (4 * m + 2 * m1 + m2)
        out["w2"] = nnz * (s + 4) + 4 * (n + 1) + s * m + s * 7 * n
        order = (["Fw", "FTw_primal"] if fused_qw else ["Qw", "qp_primal"]) + ["tempv", "dual", "w2"]
    elif Q is not None and getattr(Q, "nnz", 0) > 0:
        out["Qw"] = 0
        out["qp_primal"] = s * 12 * n + s * 2 * n1
        out["tempv"] = 0
        out["dual"] = nnz * (s + 4) + 4 * (m + 1) + s * n + s * (4 * m + 2 * m1 + m2)
        out["w2"] = nnz * (s + 4) + 4 * (n + 1) + s * m + s * 8 * n
        order = ["Qw", "qp_primal", "tempv", "dual", "w2"]
    else:
        out["primal"] = nnz * (s + 4) + 4 * (n + 1) + s * m + s * (4 * n + 2 * n1)
        out["dual"] = nnz * (s + 4) + 4 * (m + 1) + s * n + s * (3 * m + 2 * m1 + m2)
        order = ["primal", "dual"]
    return out, order


def algorithmic_bytes(model, s=8, fused_qw=False):
    m, n = model.A.shape
    return algorithmic_bytes_dims(m, n, model.A.nnz, model.number_linear_con, model.number_lu_x,
                                  model.Q, s, fused_qw)


def make_workload(name, scale):
    from hprsocp import synthetic
    if name == "c0":
        return synthetic.make_c0(seed=0), "C0 plain FP64 small SOCP (n=1000, m=610)"
    if name == "c1":
        m = synthetic.make_c1(scale=scale, seed=1000, q_form="factored")
        return m, (f"C1 large single QP-SOCP (n={m.A.shape[1]}, m_A={m.number_linear_con} + "
                   f"{len(m.SOC_con_idx) - 1} cones U[3,201], nnz(A)={m.A.nnz}, Q=F'F+D factored)")
    if name == "c2":
        m = synthetic.make_c2(scale=scale, seed=2000)
        return m, (f"C2 cone-projection stress QP-SOCP (n={m.A.shape[1]}, m={m.A.shape[0]}, "
                   f"nnz(A)={m.A.nnz}, diagonal Q)")
    raise SystemExit(f"unknown config {name}")


def oracle_lib():
    from hprsocp import _abi
    so = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(so):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True, stdout=subprocess.DEVNULL)
    olib = _abi.bind(C.CDLL(so), "oracle_")
    olib.oracle_set_threads.argtypes = [C.c_int32]
    olib.oracle_set_threads(1)
    return olib


def host_cores():
    """Host cores this process may use, capped at the 16-core share of one GPU."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(model, sigma, lam_a, lam_q, iters, threads=1):
    """The oracle (FP64 C restatement of the reference CPU path) on a bounded
    sample: `iters` HPR iterations of the same instance after setup.  `threads`
    > 1 runs its sparse products with OpenMP (the rest stays serial, as in the
    reference's single-threaded CPU path)."""
    import hprsocp
    olib = oracle_lib()
    olib.oracle_set_threads.argtypes = [C.c_int32]
    olib.oracle_set_threads(int(threads))
    p = hprsocp.HPRSOCP_parameters(verbose=False)
    with hprsocp.Solver(model, p, lib=olib, prefix="oracle_") as o:
        o.set_scalars(sigma=sigma, lambda_A=lam_a, lambda_Q=lam_q)
        o.update(0, False)
        t0 = time.perf_counter()
        for k in range(iters):
            o.update(k + 1, False)
        dt = time.perf_counter() - t0
    olib.oracle_set_threads(1)
    return iters / dt, dt


def device_info():
    """hipDeviceProp figures of device 0 (name, CUs, memory size, memory clock
    and bus width), read through the HIP runtime, to set beside the 8 TB/s
    published peak."""
    info = {}
    try:
        hip = C.CDLL("libamdhip64.so")
    except OSError:
        return info
    dev = C.c_int(0)
    name = C.create_string_buffer(256)
    if hip.hipDeviceGet(C.byref(dev), 0) == 0 and hip.hipDeviceGetName(name, 256, dev) == 0 and name.value:
        info["name"] = name.value.decode(errors="replace")
    total = C.c_size_t()
    if hip.hipDeviceTotalMem(C.byref(total), 0) == 0:
        info["mem_GiB"] = round(total.value / 2**30, 1)
    v = C.c_int()
    # hipDeviceAttribute_t: MultiprocessorCount = 63, MemoryClockRate = 60 (kHz), MemoryBusWidth = 59 (bits).
    # HBM3E moves 4 bits per pin per reported clock: 2.0 GHz x 4 x 8192 bits = 8.19 TB/s.
    for key, attr in (("cus", 63), ("mem_clock_kHz", 60), ("mem_bus_bits", 59)):
        if hip.hipDeviceGetAttribute(C.byref(v), attr, 0) == 0:
            info[key] = v.value
    return info


def elapsed():
    return time.time() - T_START


def solve_record(r, stoptol):
    hit6 = r.residuals < 1e-6 or r.iter_6 < r.iter
    ok = r.status == "OPTIMAL"
    rec = {"status": r.status, "stoptol": stoptol, "iter": r.iter, "time_s": round(r.time, 3), "kkt": r.residuals,
           "iter_to_1e-6": r.iter_6 if hit6 else None, "time_to_1e-6_s": round(r.time_6, 3) if hit6 else None,
           "primal_obj": r.primal_obj, "restarts": r.restarts, "iter_fp32": r.iter_fp32}
    if stoptol < 1e-6:
        rec[f"time_to_{stoptol:g}_s"] = round(r.time, 3) if ok else None
        rec[f"iter_to_{stoptol:g}"] = r.iter if ok else None
    return rec


def gather_measure(g, model, fp32=False, slices=8):
    """Entries per second of the iteration's two big SpMVs (A x and A^T y, in
    the form the autotuner kept; ``fp32``: the FP32 operators) next to the
    gather ceilings the probe kernel measures for the same entry count and
    slice width: random column order (the row-major layouts' access pattern)
    and column-sorted chunks."""
    m, n = model.A.shape
    nnz = int(model.A.nnz)
    out = {}
    for name, trans, width_total in (("A", 0, n), ("AT", 1, m)):
        t = g.time_spmv(trans, 20, fp32=fp32)
        width = -(-width_total // slices)
        rec = {"entries": nnz, "spmv_ms": round(t, 5), "G_per_s": round(nnz / t / 1e6, 1),
               "bytes_per_entry": 8 if fp32 else 12}
        if nnz >= 16384 and width <= (1 << 18):
            pr = g.probe_gather(nnz, width, slices, False, fp32=fp32)
            ps = g.probe_gather(nnz, width, slices, True, fp32=fp32)
            rec["ceiling_random_G_per_s"] = round(nnz / pr / 1e6, 1)
            rec["ceiling_sorted_G_per_s"] = round(nnz / ps / 1e6, 1)
            rec["frac_of_sorted_ceiling"] = round(ps / t, 3)
        out[name] = rec
    return out


def group_roofline(bytes_by, order, kern):
    """{group: ms, algorithmic MB, GB/s, fraction of the HBM peak} of the
    launch groups that ran (groups with no bytes or no time are left out)."""
    out = {}
    for i, name in enumerate(order):
        t = kern[i] if kern and i < len(kern) else 0.0
        b = bytes_by.get(name, 0)
        if t <= 0 or b <= 0:
            continue
        gbs = b / (t * 1e-3) / 1e9
        out[name] = {"ms": round(t, 5), "algorithmic_MB": round(b / 1e6, 2), "GBps": round(gbs, 1),
                     "frac": round(gbs / PEAK_HBM_GBS, 4)}
    return out


def compact_plan(info):
    """The autotuner's choice per operator, in a few fields."""
    out = {}
    for o in info.get("operators", []):
        rec = {"form": o["form"], "tuned": o["tuned"]}
        if o["tuned"]:
            rec.update({k: o[k] for k in ("nnz_target", "g_divisor", "huge_target", "huge_whole", "stream")})
            if o["candidates"]:
                rec.update({"spmv_ms_default": o["ms_default"], "spmv_ms": o["ms"], "candidates": o["candidates"]})
        if o["form"] != "plain":
            rec["spmv_ms_form"] = o["ms_form"]
        if o.get("block"):
            rec["sorted_block"] = o["block"]
        if o.get("form32"):
            rec.update({"fp32_layout": o["form32"], "fp32_block": o["block32"], "fp32_spmv_ms": o["ms32"]})
        out[o["name"]] = rec
    return out


def bench_single(args):
    import hprsocp
    hprsocp.load_library()
    t0 = time.time()
    model, wl = make_workload(args.config, args.scale)
    m, n = model.A.shape
    log(f"workload {wl} generated in {time.time() - t0:.1f}s")
    precs = precisions(args, SINGLE_PRECS)
    solve_precs = [p for p in precs if p in SOLVE_PRECS or (p == "fp32_refine" and args.refine_solve)]
    res = {}
    plan = None
    copy_gbs = None
    for prec in precs:
        p = hprsocp.HPRSOCP_parameters(verbose=False, stoptol=args.single_stoptol, time_limit=args.solve_time,
                                       precision=prec, use_graph=not args.no_graph)
        t0 = time.time()
        g = hprsocp.Solver(model, p)
        log(f"[{prec}] setup (transfer + scaling + plans) {time.time() - t0:.2f}s")
        sc = g.estimate_lambda()
        log(f"[{prec}] lambda_A={sc['lambda_max_A']:.6e} lambda_Q={sc['lambda_max_Q']:.6e} sigma={sc['sigma']:.4e}")
        if plan is None:
            plan = g.plan_info()
            # a plain device copy: the streaming bandwidth the kernels can reach on this box
            cmb = args.probe_copy_mb or 512
            pms = g.probe_copy(cmb << 20)
            copy_gbs = 2 * (cmb << 20) / (pms * 1e-3) / 1e9
            log(f"copy probe: {cmb} MiB in {pms:.4f} ms ({copy_gbs:.0f} GB/s read+write)")
        g.time_iterations(args.warmup)
        ms, _ = g.time_iterations(args.steps)
        _, kern = g.time_iterations(max(args.steps // 4, 10), per_kernel=True)
        log(f"[{prec}] {ms:.4f} ms/iter -> {1000.0 / ms:.1f} iters/s; per-group ms {kern}")
        gather = None
        if not args.no_gather and prec in ("fp64", "fp32_delta"):
            gather = gather_measure(g, model, fp32=prec != "fp64")
            log(f"[{prec}] SpMV gathers vs probe ceilings: {gather}")
        g.close()
        rec = {"ms_per_iter": ms, "kern": kern, "sc": sc, "solve": None, "gather": gather}
        if not args.no_solve and prec in solve_precs:
            t0 = time.time()
            with hprsocp.Solver(model, p) as g2:
                g2.estimate_lambda()
                r = g2.solve()
            rec["solve"] = solve_record(r, args.single_stoptol)
            log(f"[{prec}] solve: {rec['solve']} (wall {time.time() - t0:.1f}s)")
        res[prec] = rec

    base = res[precs[0]]
    ms = base["ms_per_iter"]
    vbytes = 8 if precs[0] == "fp64" else 4
    fused_qw = bool(plan and plan.get("qw_primal_fused"))
    bytes_by, order = algorithmic_bytes(model, vbytes, fused_qw)
    groups = [(name, base["kern"][i]) for i, name in enumerate(order)]
    dom, dom_ms = max(groups, key=lambda t: t[1])
    dom_bytes = bytes_by[dom]
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    it_bytes = sum(bytes_by.values())
    a_form = (compact_plan(plan).get(GROUP_OPERATOR.get(dom, "A"), {}).get("form")
              if plan else None)
    traffic = (MEASURED_TRAFFIC.get((args.config, precs[0], dom, a_form, KERNEL_REV))
               if args.scale == 1.0 else None)

    ncores = host_cores()
    if args.cpu_iters > 0:
        sc = base["sc"]
        cpu_rate, cpu_dt = cpu_baseline(model, sc["sigma"], sc["lambda_max_A"], sc["lambda_max_Q"], args.cpu_iters)
        log(f"cpu oracle: {cpu_rate:.3f} iters/s over {args.cpu_iters} iterations ({cpu_dt:.1f}s)")
        mt_rate, mt_dt = cpu_baseline(model, sc["sigma"], sc["lambda_max_A"], sc["lambda_max_Q"], args.cpu_iters,
                                      threads=ncores)
        log(f"cpu oracle, {ncores} threads: {mt_rate:.3f} iters/s ({mt_dt:.1f}s)")
    else:
        cpu_rate, cpu_dt, mt_rate, mt_dt = None, 0.0, None, 0.0

    forms = {}
    for prec, rec in res.items():
        vb = 8 if prec == "fp64" else 4
        bb, od = algorithmic_bytes(model, vb, fused_qw)
        forms[prec] = {"iters_per_s": round(1000.0 / rec["ms_per_iter"], 2),
                       "ms_per_iter": round(rec["ms_per_iter"], 5),
                       "groups": group_roofline(bb, od, rec["kern"]),
                       "groups_value_bytes": vb, "solve": rec["solve"]}
        if rec["gather"]:
            forms[prec]["gather"] = rec["gather"]
    if "fp64" in res:
        vs = {}
        for prec in ("fp32_delta", "fp32_refine"):
            if prec not in res:
                continue
            s64, s32 = res["fp64"]["solve"], res[prec]["solve"]
            vs[prec] = {"iter_rate": round(res["fp64"]["ms_per_iter"] / res[prec]["ms_per_iter"], 3)}
            for key in ("time_to_1e-6_s", f"time_to_{args.single_stoptol:g}_s"):
                if s64 and s32 and s64.get(key) and s32.get(key):
                    vs[prec][key.replace("_s", "").replace("time_to_", "time_to_")] = round(s64[key] / s32[key], 3)
        if vs:
            forms["vs_fp64"] = vs

    return {
        "metric": METRIC,
        "value": round(1000.0 / ms, 3),
        "unit": "HPR iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if precs[0] == "fp64" else "f32",
        "data": "synthetic (seeded feasibility construction)",
        "device": device_info(),
        "config": {"workload": wl, "config": args.config, "scale": args.scale, "m": m, "n": n,
                   "nnz_A": int(model.A.nnz), "precision": precs[0],
                   "algorithmic_MB_per_iter": round(it_bytes / 1e6, 2),
                   "iteration_GBps": round(it_bytes / (ms * 1e-3) / 1e9, 1),
                   "plan": compact_plan(plan) if plan else None, "qw_primal_fused": fused_qw},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                     "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
                     "traffic": traffic, "traffic_key": [args.config, precs[0], dom, a_form, KERNEL_REV],
                     "algorithmic_bytes": int(dom_bytes), "ms": round(dom_ms, 5),
                     "gather": base["gather"], "groups": forms[precs[0]]["groups"],
                     "copy_GBps": round(copy_gbs, 1) if copy_gbs else None},
        "kernel_ms": {name: round(t, 5) for name, t in groups},
        "forms": forms,
        "cpu_baseline": {"value": cpu_rate and round(cpu_rate, 4), "unit": "HPR iters/s", "cores": 1,
                         "kind": "port",
                         "sample": f"{args.cpu_iters} partial HPR iterations of the same instance "
                                   f"after setup (FP64 C oracle, 1 thread, {cpu_dt:.1f}s)",
                         "all_cores": {"value": mt_rate and round(mt_rate, 4), "cores": ncores,
                                       "sample": f"the same iterations with OpenMP sparse products "
                                                 f"({mt_dt:.1f}s)"}},
    }


def batch_workload(name, count):
    from hprsocp import synthetic
    if name == "c3":
        models = synthetic.make_c3(count=count or 512, seed=3000)
        warms = [None] * len(models)
        desc = "portfolio SOCPs (n log-uniform [200, 20000], k in [5, 50])"
    else:
        models, warms = synthetic.make_c4(count=count or 64, seed=4000)
        desc = ("robust-regression QP-SOCPs (d ~ U[1e4, 2e5], diagonal Q, one SOC of size d+1; "
                "even-indexed half warm-started from perturbed solutions)")
    return models, warms, desc


def batch_bytes(models, s=8):
    """Algorithmic bytes per partial iteration of the packed batch, per launch group."""
    m = sum(mm.A.shape[0] for mm in models)
    n = sum(mm.A.shape[1] for mm in models)
    nnz = sum(mm.A.nnz for mm in models)
    m1 = sum(mm.number_linear_con for mm in models)
    n1 = sum(mm.number_lu_x for mm in models)
    return algorithmic_bytes_dims(m, n, nnz, m1, n1, models[0].Q, s)


def bench_batch(args, name="c3"):
    """Batched solves vs sequential single-problem FP64 solves of every instance
    (C3 portfolios, C4 regressions), at each tolerance of the config."""
    import numpy as np
    import hprsocp
    hprsocp.load_library()
    t0 = time.time()
    models, warms, desc = batch_workload(name, args.batch_count if name == "c3" else args.c4_count)
    N = len(models)
    sum_n = sum(mm.A.shape[1] for mm in models)
    sum_m = sum(mm.A.shape[0] for mm in models)
    sum_nnz = sum(mm.A.nnz for mm in models)
    wl = f"{name.upper()} batch of {N} {desc}; sum n={sum_n}, sum m={sum_m}, sum nnz={sum_nnz}"
    log(f"workload {wl} generated in {time.time() - t0:.1f}s")
    tols = [args.stoptol] if name == "c3" else sorted({args.stoptol, 1e-8}, reverse=True)
    bprecs = precisions(args, BATCH_PRECS if name == "c3" else C4_PRECS)
    out = {"workload": wl, "count": N, "tolerances": tols}
    for tol in tols:
        key = f"{tol:g}"
        rt = {}
        objs = None
        for prec in bprecs:
            p = hprsocp.HPRSOCP_parameters(verbose=False, stoptol=tol, time_limit=args.solve_time, precision=prec,
                                           use_graph=not args.no_graph)
            t0 = time.time()
            b = hprsocp.BatchSolver(models, p)
            b.estimate_lambda()
            if any(w is not None for w in warms):
                b.set_initial(warms, None)
            t_setup = time.time() - t0
            t1 = time.time()
            rs = b.solve()
            t_solve = time.time() - t1
            wall = time.time() - t0
            rec = {"solves_per_s": round(N / wall, 3), "wall_s": round(wall, 3), "setup_s": round(t_setup, 3),
                   "solve_s": round(t_solve, 3), "status": {}, "max_iter": max(r.iter for r in rs),
                   "median_iter": int(np.median([r.iter for r in rs])), "max_kkt": max(r.residuals for r in rs)}
            for r in rs:
                rec["status"][r.status] = rec["status"].get(r.status, 0) + 1
            if tol == tols[0]:
                ms_it, kern = b.time_iterations(max(args.steps // 4, 20), per_kernel=True)
                vb = 8 if prec == "fp64" else 4
                bb, od = batch_bytes(models, vb)
                rec["ms_per_iter"] = round(ms_it, 5)
                rec["groups"] = group_roofline(bb, od, kern)
                rec["groups_value_bytes"] = vb
                if prec == bprecs[0]:
                    out["plan"] = compact_plan(b.plan_info())
            b.close()
            rec["control"] = "device"
            if prec == "fp64" and tol == tols[0]:
                # the same solve with the window-edge steps on the host, for comparison
                saved = os.environ.get("HPRSOCP_DEVICE_CONTROL")
                os.environ["HPRSOCP_DEVICE_CONTROL"] = "0"
                try:
                    t0 = time.time()
                    with hprsocp.BatchSolver(models, p) as bh:
                        bh.estimate_lambda()
                        if any(w is not None for w in warms):
                            bh.set_initial(warms, None)
                        t1 = time.time()
                        rh = bh.solve()
                        t2 = time.time()
                finally:
                    if saved is None:
                        del os.environ["HPRSOCP_DEVICE_CONTROL"]
                    else:
                        os.environ["HPRSOCP_DEVICE_CONTROL"] = saved
                rec["host_control"] = {"solves_per_s": round(N / (t2 - t0), 3), "solve_s": round(t2 - t1, 3),
                                       "same_iterations": [r.iter for r in rh] == [r.iter for r in rs]}
            rt[prec] = rec
            log(f"[{name} batch {prec} @{key}] {rec}")
            if prec == "fp64":
                objs = [r.primal_obj for r in rs]
        # every instance solved sequentially by the single-problem FP64 path
        if elapsed() < BUDGET_S:
            p = hprsocp.HPRSOCP_parameters(verbose=False, stoptol=tol, time_limit=args.solve_time,
                                           use_graph=not args.no_graph)
            idx = range(N) if args.seq_sample <= 0 else np.linspace(0, N - 1, min(args.seq_sample, N)).round()
            idx = [int(i) for i in idx]
            t0 = time.time()
            worst = 0.0
            for i in idx:
                with hprsocp.Solver(models[i], p) as g:
                    g.estimate_lambda()
                    if warms[i] is not None:
                        g.set_initial(warms[i], None)
                    r = g.solve()
                if objs is not None:
                    worst = max(worst, abs(r.primal_obj - objs[i]) / (1 + abs(r.primal_obj)))
            t_seq = time.time() - t0
            seq = {"solves": len(idx), "wall_s": round(t_seq, 3), "solves_per_s": round(len(idx) / t_seq, 3),
                   "all_instances": len(idx) == N, "batch_vs_single_max_obj_rel_diff": worst}
            rt["sequential_fp64"] = seq
            log(f"[{name} sequential fp64 @{key}] {seq}")
            if len(idx) == N:
                rt["batch_vs_sequential"] = {prec: round(rt[prec]["solves_per_s"] / seq["solves_per_s"], 2)
                                             for prec in bprecs}
        else:
            rt["sequential_fp64"] = {"skipped": f"run time budget ({BUDGET_S:.0f} s) reached"}
        out[key] = rt
    if args.cpu_iters > 0 and elapsed() < BUDGET_S:
        # CPU oracle (1 thread) on an evenly spaced sample of the instances (C3: 8;
        # C4: the 2 smallest, its instances take the oracle tens of seconds each),
        # next to the GPU's sequential FP64 solves of the same sample
        olib = oracle_lib()
        if name == "c3":
            sample = [int(i) for i in np.linspace(0, N - 1, min(8, N)).round()]
            what = f"{len(sample)} evenly spaced instances"
        else:
            sample = sorted(range(N), key=lambda i: models[i].A.nnz)[:2]
            what = f"the {len(sample)} smallest instances (by nnz)"
        p = hprsocp.HPRSOCP_parameters(verbose=False, stoptol=tols[0], time_limit=args.solve_time)
        t_cpu = t_gpu = 0.0
        for i in sample:
            for lib, pref in ((olib, "oracle_"), (None, "hprsocp_")):
                t0 = time.time()
                with hprsocp.Solver(models[i], p, lib=lib, prefix=pref) as o:
                    o.estimate_lambda()
                    if warms[i] is not None:
                        o.set_initial(warms[i], None)
                    o.solve()
                if lib is None:
                    t_gpu += time.time() - t0
                else:
                    t_cpu += time.time() - t0
        out["cpu_oracle"] = {"sample": f"{what}, FP64 C oracle, 1 thread; no extrapolation", "stoptol": tols[0],
                             "wall_s": round(t_cpu, 3), "solves_per_s_of_sample": round(len(sample) / t_cpu, 4),
                             "gpu_sequential_fp64_solves_per_s_of_sample": round(len(sample) / t_gpu, 3)}
        log(f"[{name} cpu oracle] {out['cpu_oracle']}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", default="c1")
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--solve-time", type=float, default=90.0, help="time cap of each solve")
    ap.add_argument("--cpu-iters", type=int, default=100,
                    help="C1 iterations of the 1-thread CPU oracle sample (about 12 s; also run on all cores)")
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched C3 and C4 measurements")
    ap.add_argument("--no-c4", action="store_true", help="skip the batched C4 measurement")
    ap.add_argument("--no-graph", action="store_true", help="launch every update eagerly (no HIP graphs)")
    ap.add_argument("--no-gather", action="store_true", help="skip the SpMV gather-ceiling measurement")
    ap.add_argument("--refine-solve", action="store_true", help="also solve in fp32_refine")
    ap.add_argument("--precision", default="all", choices=["fp64", "fp32_refine", "fp32_delta", "both", "all"],
                    help="all: fp64, fp32_delta and fp32_refine; both: fp64 and fp32_delta")
    ap.add_argument("--batch-count", type=int, default=512)
    ap.add_argument("--seq-sample", type=int, default=0, help="sequential solves per batch (0: every instance)")
    ap.add_argument("--c4-count", type=int, default=64)
    ap.add_argument("--stoptol", type=float, default=1e-6, help="stoptol of the batched configs (C4 adds 1e-8)")
    ap.add_argument("--single-stoptol", type=float, default=1e-8,
                    help="stoptol of the single-problem solves (the 1e-6 milestone is read on the way)")
    ap.add_argument("--probe-copy-mb", type=int, default=0,
                    help="also run one plain device copy of this many MiB (byte-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the HPR path here is "
              "single-GPU — see the sharding design notes.", file=sys.stderr)
        sys.exit(2)
    if args.config in ("c3", "c4"):
        b = bench_batch(args, args.config)
        t0 = b[f"{b['tolerances'][0]:g}"]
        present = [k for k in BATCH_PRECS if k in t0]
        first = t0[present[0]]
        best = max(t0[k]["solves_per_s"] for k in present)
        out = {"metric": METRIC, "value": best, "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
               "warmup": args.warmup, "ms_per_step": first["ms_per_iter"], "higher_is_better": True,
               "scaling": "weak", "vs_baseline": None, "dtype": "f64" if present[0] == "fp64" else "f32",
               "data": "synthetic (seeded)",
               "config": {"workload": b["workload"], "config": args.config, "stoptol": args.stoptol}, "batch": b,
               "cpu_baseline": {"value": b.get("cpu_oracle", {}).get("solves_per_s_of_sample"), "unit": "solves/s",
                                "cores": 1, "kind": "port",
                                "sample": b.get("cpu_oracle", {}).get("sample")}}
        print(json.dumps(out), flush=True)
        return
    out = bench_single(args)
    if not args.no_batch and args.config == "c1":
        for name in ("c3", "c4"):
            if name == "c4" and args.no_c4:
                continue
            if elapsed() > BUDGET_S:
                out[f"batch_{name}"] = {"skipped": f"run time budget ({BUDGET_S:.0f} s) reached"}
                continue
            try:
                out[f"batch_{name}"] = bench_batch(args, name)
            except Exception as e:  # the single-problem line stands on its own
                log(f"{name} batch measurement failed: {e!r}")
                out[f"batch_{name}"] = {"error": repr(e)}
    out["summary"] = headline_summary(out, args)
    out["bench_wall_s"] = round(elapsed(), 1)
    print(json.dumps(out), flush=True)


def headline_summary(out, args):
    """The headline comparisons in one place: FP64 vs the FP32 delta form in
    time to 1e-6 and to the single-problem tolerance, and each batch against
    the sequential FP64 solves of all its instances (same-precision FP64 ratio
    and the best form's ratio)."""
    s = {"fp64_iters_per_s": out.get("value")}
    forms = out.get("forms", {})
    s64 = (forms.get("fp64") or {}).get("solve") or {}
    sdl = (forms.get("fp32_delta") or {}).get("solve") or {}
    for key in ("time_to_1e-6_s", f"time_to_{args.single_stoptol:g}_s"):
        if s64.get(key) and sdl.get(key):
            tag = key.replace("time_to_", "").replace("_s", "")
            s[f"fp64_time_to_{tag}_s"] = s64[key]
            s[f"fp32_delta_time_to_{tag}_s"] = sdl[key]
            s[f"fp32_delta_vs_fp64_time_to_{tag}"] = round(s64[key] / sdl[key], 3)
    for name in ("c3", "c4"):
        b = out.get(f"batch_{name}") or {}
        for tol in b.get("tolerances", []):
            rt = b.get(f"{tol:g}") or {}
            ratios = rt.get("batch_vs_sequential")
            if not ratios:
                continue
            best = max(ratios, key=ratios.get)
            s[f"{name}@{tol:g}"] = {"sequential_fp64_solves_per_s": rt["sequential_fp64"]["solves_per_s"],
                                    "fp64_batch_solves_per_s": rt["fp64"]["solves_per_s"],
                                    "fp64_batch_vs_sequential": ratios.get("fp64"),
                                    "best_form": best, "best_batch_solves_per_s": rt[best]["solves_per_s"],
                                    "best_vs_sequential": ratios[best]}
    return s


if __name__ == "__main__":
    main()
