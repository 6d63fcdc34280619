#!/usr/bin/env python3
"""Benchmark of the HPR iteration on one MI355X.

Single problem (default --config c1): the large QP-SOCP, n = 1e6 boxed
variables, m_A = 5e5 linear rows with log-normal row lengths (mean 20) plus
10,000 constraint cones of sizes U[3,201], Q = F'F + diag(U(0,1)) with F
1e5 x 1e6 at 5 nnz per column (factored operator form); seeded synthetic data.
A "step" is one HPR iteration (main update, partial outputs) over the resident
instance; `value` = FP64 HPR iterations/s.  Also measured: FP32+refine
iterations/s of the two FP32 forms (fp32_refine: FP32 state, FP64 checks, FP64
takeover; fp32_delta: FP32 differences from the FP64 restart anchors), time to
KKT 1e-6 of full solves in all three precisions (time-capped),
the roofline of the dominant launch group (HIP events on the solver's stream)
and the CPU oracle on a bounded sample.

Batch (--config c3, and part of the default run unless --no-batch): 512
portfolio SOCPs, n log-uniform in [200, 20000], solved together by the batched
backend (all three precisions) against sequential single-problem FP64 solves
(timed on an evenly spaced sample and scaled to the batch).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c1|c2|c0|c3] [--scale s]
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
SINGLE_PRECS = ["fp64", "fp32_refine", "fp32_delta"]
BATCH_PRECS = ["fp64", "fp32_refine", "fp32_delta"]


def precisions(args, every):
    if args.precision == "all":
        return every
    return ["fp64", "fp32_refine"] if args.precision == "both" else [args.precision]
PEAK_HBM_GBS = 8000.0

# HBM bytes per launch of the dominant launch group, from two rocprofv3 --pmc
# passes (FETCH_SIZE, WRITE_SIZE) over this script, each scaled by what the same
# counter reads for the 1 GiB plain copy of hprsocp_probe_copy (read factor 2.0,
# write factor 1.0).  Source: profiles/c1_fp64_iteration_hbm_traffic.txt.
# Keyed by (config, precision, group); only full-size runs use it.
MEASURED_TRAFFIC = {
    ("c1", "fp64", "dual"): 537.4e6 + 172.4e6,   # k_spmv_sorted(A, 8 slices) + k_tile<OpDual, slices>
}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes_dims(m, n, nnz, m1, n1, Q, s=8):
    """Compulsory HBM bytes per partial HPR iteration, per launch group (s = value bytes).

    Each matrix is read once per SpMV (s + 4 bytes per nonzero, 4 bytes per row
    pointer), each vector once per kernel; gathered SpMV inputs count once."""
    m2 = m - m1
    out = {}
    if isinstance(Q, tuple):
        F = Q[1]
        mf, nnzf = F.shape[0], F.nnz
        f1 = nnzf * (s + 4) + 4 * (mf + 1) + s * (n + mf)             # F v
        f2 = nnzf * (s + 4) + 4 * (n + 1) + s * (mf + 3 * n)          # F'(.) + D v
        out["Qw"] = f1 + f2
        out["qp_primal"] = s * (8 * n + 3 * n)
        out["tempv"] = f1 + f2 + s * 3 * n                            # + x_hat, qw, tempv
        out["dual"] = nnz * (s + 4) + 4 * (m + 1) + s * n + s * (4 * m + 2 * m1 + m2)
        out["w2"] = nnz * (s + 4) + 4 * (n + 1) + s * m + s * 7 * n
        order = ["Qw", "qp_primal", "tempv", "dual", "w2"]
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


def algorithmic_bytes(model, s=8):
    m, n = model.A.shape
    return algorithmic_bytes_dims(m, n, model.A.nnz, model.number_linear_con, model.number_lu_x,
                                  model.Q, s)


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


def solve_record(r, stoptol=1e-6):
    hit6 = r.residuals < 1e-6 or r.iter_6 < r.iter
    rec = {"status": r.status, "iter": r.iter, "time_s": round(r.time, 3), "kkt": r.residuals,
           "iter_to_1e-6": r.iter_6 if hit6 else None, "time_to_1e-6_s": round(r.time_6, 3) if hit6 else None,
           "primal_obj": r.primal_obj, "restarts": r.restarts, "iter_fp32": r.iter_fp32}
    if stoptol < 1e-6:
        ok = r.status == "OPTIMAL"
        rec[f"time_to_{stoptol:g}_s"] = round(r.time, 3) if ok else None
        rec[f"iter_to_{stoptol:g}"] = r.iter if ok else None
    return rec


def gather_measure(g, model, slices=8):
    """Gathers per second of the iteration's two big SpMVs (A x and A^T y, as
    the autotuner left them) next to the gather ceilings the probe kernel
    measures for the same entry count and slice width: random column order
    (the access pattern of the row-major layouts) and column-sorted chunks."""
    m, n = model.A.shape
    nnz = int(model.A.nnz)
    out = {}
    for name, trans, width_total in (("A", 0, n), ("AT", 1, m)):
        t = g.time_spmv(trans, 20)
        width = -(-width_total // slices)
        rec = {"entries": nnz, "spmv_ms": round(t, 5), "G_per_s": round(nnz / t / 1e6, 1)}
        if nnz >= 8192 and width <= (1 << 19):
            pr = g.probe_gather(nnz, width, slices, False)
            ps = g.probe_gather(nnz, width, slices, True)
            rec["ceiling_random_G_per_s"] = round(nnz / pr / 1e6, 1)
            rec["ceiling_sorted_G_per_s"] = round(nnz / ps / 1e6, 1)
        out[name] = rec
    return out


def bench_single(args):
    import hprsocp
    hprsocp.load_library()
    t0 = time.time()
    model, wl = make_workload(args.config, args.scale)
    m, n = model.A.shape
    log(f"workload {wl} generated in {time.time() - t0:.1f}s")
    precs = precisions(args, SINGLE_PRECS)
    res = {}
    for prec in precs:
        p = hprsocp.HPRSOCP_parameters(verbose=False, stoptol=args.single_stoptol, time_limit=args.solve_time,
                                       precision=prec, use_graph=not args.no_graph)
        t0 = time.time()
        g = hprsocp.Solver(model, p)
        log(f"[{prec}] setup (transfer + scaling + plans) {time.time() - t0:.2f}s")
        sc = g.estimate_lambda()
        log(f"[{prec}] lambda_A={sc['lambda_max_A']:.6e} lambda_Q={sc['lambda_max_Q']:.6e} sigma={sc['sigma']:.4e}")
        if prec == precs[0]:
            # a plain device copy: the streaming bandwidth the kernels can reach on this box
            cmb = args.probe_copy_mb or 512
            pms = g.probe_copy(cmb << 20)
            copy_gbs = 2 * (cmb << 20) / (pms * 1e-3) / 1e9
            log(f"copy probe: {cmb} MiB in {pms:.4f} ms ({copy_gbs:.0f} GB/s read+write)")
        g.time_iterations(args.warmup)
        ms, _ = g.time_iterations(args.steps)
        _, kern = g.time_iterations(max(args.steps // 4, 10), per_kernel=True)
        log(f"[{prec}] {ms:.4f} ms/iter -> {1000.0 / ms:.1f} iters/s; per-group ms {kern}")
        gather = gather_measure(g, model) if prec == precs[0] and not args.no_gather else None
        if gather:
            log(f"[{prec}] SpMV gathers vs probe ceilings: {gather}")
        g.close()
        rec = {"ms_per_iter": ms, "kern": kern, "sc": sc, "solve": None, "gather": gather}
        if not args.no_solve:
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
    bytes_by, order = algorithmic_bytes(model, vbytes)
    groups = [(name, base["kern"][i]) for i, name in enumerate(order)]
    group_roof = {name: {"ms": round(t, 5), "algorithmic_MB": round(bytes_by[name] / 1e6, 2),
                         "GBps": round(bytes_by[name] / (t * 1e-3) / 1e9, 1) if t > 0 else None,
                         "frac": round(bytes_by[name] / (t * 1e-3) / 1e9 / PEAK_HBM_GBS, 4) if t > 0 else None}
                  for name, t in groups}
    dom, dom_ms = max(groups, key=lambda t: t[1])
    dom_bytes = bytes_by[dom]
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    it_bytes = sum(bytes_by.values())
    traffic = MEASURED_TRAFFIC.get((args.config, precs[0], dom)) if args.scale == 1.0 else None

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
        forms[prec] = {"iters_per_s": round(1000.0 / rec["ms_per_iter"], 2),
                       "ms_per_iter": round(rec["ms_per_iter"], 5), "solve_to_1e-6": rec["solve"]}
    if "fp64" in res:
        vs = {}
        for prec in ("fp32_refine", "fp32_delta"):
            if prec not in res:
                continue
            s64, s32 = res["fp64"]["solve"], res[prec]["solve"]
            vs[prec] = {"iter_rate": round(res["fp64"]["ms_per_iter"] / res[prec]["ms_per_iter"], 3)}
            if s64 and s32 and s64["time_to_1e-6_s"] and s32["time_to_1e-6_s"]:
                vs[prec]["time_to_1e-6"] = round(s64["time_to_1e-6_s"] / s32["time_to_1e-6_s"], 3)
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
                   "iteration_GBps": round(it_bytes / (ms * 1e-3) / 1e9, 1)},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                     "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
                     "traffic": traffic, "algorithmic_bytes": int(dom_bytes), "ms": round(dom_ms, 5),
                     "gather": base["gather"], "groups": group_roof,
                     "copy_GBps": round(copy_gbs, 1)},
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


def bench_batch(args, name="c3"):
    """Batched solves vs sequential single-problem FP64 solves (C3 portfolios, C4 regressions)."""
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
    tol = args.stoptol
    out = {"workload": wl, "count": N}
    objs = None
    bprecs = precisions(args, BATCH_PRECS)
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
        ms_it, kern = b.time_iterations(max(args.steps // 4, 20), per_kernel=True)
        b.close()
        st = {}
        for r in rs:
            st[r.status] = st.get(r.status, 0) + 1
        rec = {"solves_per_s": round(N / wall, 3), "wall_s": round(wall, 3), "setup_s": round(t_setup, 3),
               "solve_s": round(t_solve, 3), "status": st, "max_iter": max(r.iter for r in rs),
               "median_iter": int(np.median([r.iter for r in rs])), "ms_per_iter": round(ms_it, 5),
               "max_kkt": max(r.residuals for r in rs), "kernel_ms": [round(k, 5) for k in kern[:2]]}
        out[prec] = rec
        log(f"[batch {prec}] {rec}")
        if prec == "fp64":
            objs = [r.primal_obj for r in rs]
    # sequential single-problem FP64 solves on an evenly spaced sample
    K = min(args.seq_sample if name == "c3" else args.c4_seq_sample, N)
    idx = np.linspace(0, N - 1, K).round().astype(int)
    p = hprsocp.HPRSOCP_parameters(verbose=False, stoptol=tol, time_limit=args.solve_time,
                                   use_graph=not args.no_graph)
    t0 = time.time()
    worst = 0.0
    for i in idx:
        with hprsocp.Solver(models[i], p) as g:
            g.estimate_lambda()
            if warms[i] is not None:
                g.set_initial(warms[i], None)
            r = g.solve()
            g.get("x_bar")
        if objs is not None:
            worst = max(worst, abs(r.primal_obj - objs[i]) / (1 + abs(r.primal_obj)))
    t_seq = time.time() - t0
    frac = sum(models[i].A.nnz for i in idx) / sum_nnz
    est = t_seq / frac  # scaled by the sample's share of the nonzeros
    out["sequential_fp64"] = {"sample": int(K), "sample_wall_s": round(t_seq, 3),
                              "estimated_wall_s": round(est, 3), "solves_per_s": round(N / est, 3),
                              "batch_vs_single_max_obj_rel_diff": worst}
    log(f"[sequential fp64] {out['sequential_fp64']}")
    if args.cpu_iters > 0:  # CPU oracle on part of the same sample (1 thread)
        olib = oracle_lib()
        Kc = min(8 if name == "c3" else 2, K)
        t0 = time.time()
        for i in idx[:Kc]:
            with hprsocp.Solver(models[i], p, lib=olib, prefix="oracle_") as o:
                o.estimate_lambda()
                if warms[i] is not None:
                    o.set_initial(warms[i], None)
                o.solve()
        t_cpu = time.time() - t0
        frac_c = sum(models[i].A.nnz for i in idx[:Kc]) / sum_nnz
        out["cpu_oracle"] = {"sample": int(Kc), "sample_wall_s": round(t_cpu, 3),
                             "solves_per_s": round(N / (t_cpu / frac_c), 4)}
        log(f"[cpu oracle] {out['cpu_oracle']}")
    best = max(out[k]["solves_per_s"] for k in BATCH_PRECS if k in out)
    out["batch_vs_sequential"] = round(best / out["sequential_fp64"]["solves_per_s"], 2)
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
    ap.add_argument("--no-batch", action="store_true", help="skip the batched C3 measurement")
    ap.add_argument("--no-graph", action="store_true", help="launch every update eagerly (no HIP graphs)")
    ap.add_argument("--no-gather", action="store_true", help="skip the SpMV gather-ceiling measurement")
    ap.add_argument("--precision", default="all", choices=["fp64", "fp32_refine", "fp32_delta", "both", "all"],
                    help="all: fp64, fp32_refine and fp32_delta; both: fp64 and fp32_refine")
    ap.add_argument("--batch-count", type=int, default=512)
    ap.add_argument("--seq-sample", type=int, default=32)
    ap.add_argument("--c4-count", type=int, default=64)
    ap.add_argument("--c4-seq-sample", type=int, default=8)
    ap.add_argument("--stoptol", type=float, default=1e-6, help="stoptol of the batched configs")
    ap.add_argument("--single-stoptol", type=float, default=1e-6, help="stoptol of the single-problem solves")
    ap.add_argument("--probe-copy-mb", type=int, default=0,
                    help="also run one plain device copy of this many MiB (byte-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the HPR path here is "
              "single-GPU — see the sharding design notes.", file=sys.stderr)
        sys.exit(2)
    if args.config in ("c3", "c4"):
        b = bench_batch(args, args.config)
        present = [k for k in BATCH_PRECS if k in b]
        first = b[present[0]]
        best = max(b[k]["solves_per_s"] for k in present)
        out = {"metric": METRIC, "value": best, "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
               "warmup": args.warmup, "ms_per_step": first["ms_per_iter"], "higher_is_better": True,
               "scaling": "weak", "vs_baseline": None, "dtype": "f64" if present[0] == "fp64" else "f32",
               "data": "synthetic (seeded)",
               "config": {"workload": b["workload"], "config": args.config, "stoptol": args.stoptol}, "batch": b,
               "cpu_baseline": {"value": b.get("cpu_oracle", {}).get("solves_per_s"), "unit": "solves/s",
                                "cores": 1, "kind": "port",
                                "sample": "sequential oracle solves of an evenly spaced sample, scaled by nnz"}}
        print(json.dumps(out), flush=True)
        return
    out = bench_single(args)
    if not args.no_batch and args.config == "c1":
        try:
            out["batch_c3"] = bench_batch(args)
        except Exception as e:  # the single-problem line stands on its own
            log(f"batch measurement failed: {e!r}")
            out["batch_c3"] = {"error": repr(e)}
    out["summary"] = headline_summary(out)
    print(json.dumps(out), flush=True)


def headline_summary(out):
    """The headline comparisons in one place: FP64 vs the best FP32 form in time
    to the single-problem tolerance, and the best batched form vs sequential."""
    s = {"fp64_iters_per_s": out.get("value")}
    forms = out.get("forms", {})
    t64 = ((forms.get("fp64") or {}).get("solve_to_1e-6") or {}).get("time_to_1e-6_s")
    best = None
    for prec in ("fp32_refine", "fp32_delta"):
        t = ((forms.get(prec) or {}).get("solve_to_1e-6") or {}).get("time_to_1e-6_s")
        if t and (best is None or t < best[1]):
            best = (prec, t)
    if t64 and best:
        s.update({"fp64_time_to_1e-6_s": t64, "best_fp32_form": best[0], "best_fp32_time_to_1e-6_s": best[1],
                  "fp32_vs_fp64_time_to_1e-6": round(t64 / best[1], 3)})
    b = out.get("batch_c3") or {}
    if "batch_vs_sequential" in b:
        s["c3_best_batch_solves_per_s"] = max(b[k]["solves_per_s"] for k in BATCH_PRECS if k in b)
        s["c3_batch_vs_sequential"] = b["batch_vs_sequential"]
    return s


if __name__ == "__main__":
    main()
