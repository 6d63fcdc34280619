#!/usr/bin/env python3
"""Benchmark of the HPR iteration on one MI355X.

Workload (default): the large single QP-SOCP — n = 1e6 boxed variables,
m_A = 5e5 linear rows with log-normal row lengths (mean 20) plus 10,000
constraint cones of sizes U[3,201], Q = F'F + diag(U(0,1)) with F 1e5 x 1e6 at
5 nnz per column (factored operator form), seeded synthetic data.

A "step" is one HPR iteration (main update, partial outputs) over the resident
instance.  `value` = HPR iterations/s (FP64).  Also reported: time to KKT 1e-6
of a full solve (with a time cap), the roofline of the dominant kernel (HIP
events on the solver's stream) and the CPU oracle timed on a bounded sample.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c1|c2|c0] [--scale s]
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
PEAK_HBM_GBS = 8000.0

# HBM bytes per launch of the dominant launch group, from two rocprofv3 --pmc
# passes (FETCH_SIZE, WRITE_SIZE) over this script, each scaled by what the same
# counter reads for the 1 GiB plain copy of hprsocp_probe_copy (read factor 2.0,
# write factor 1.0).  Source: profiles/c1_fp64_iteration_hbm_traffic.txt.
# Keyed by (config, precision, group); only full-size runs use it.
MEASURED_TRAFFIC = {
    ("c1", "fp64", "dual"): 549.1e6 + 172.3e6,   # k_spmv_slices(A) + k_tile<OpDual, slices>
}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(model, s=8):
    """Compulsory HBM bytes per partial HPR iteration, per launch group (s = value bytes).

    Each matrix is read once per SpMV (s + 4 bytes per nonzero, 4 bytes per row
    pointer), each vector once per kernel; gathered SpMV inputs count once."""
    A = model.A
    m, n = A.shape
    nnz = A.nnz
    m1 = model.number_linear_con
    m2 = m - m1
    n1 = model.number_lu_x
    Q = model.Q
    out = {}
    if isinstance(Q, tuple):
        F = Q[1]
        mf, nnzf = F.shape[0], F.nnz
        f1 = nnzf * (s + 4) + 4 * (mf + 1) + s * (n + mf)             # F v
        f2 = nnzf * (s + 4) + 4 * (n + 1) + s * (mf + 3 * n)          # F'(.) + D v
        out["Qw"] = f1 + f2
        out["qp_primal"] = s * (8 * n + 3 * n) + s * 0 * n1
        out["tempv"] = f1 + f2 + s * 3 * n                            # + x_hat, qw, tempv
        out["dual"] = nnz * (s + 4) + 4 * (m + 1) + s * n + s * (4 * m + 2 * m1 + m2)
        out["w2"] = nnz * (s + 4) + 4 * (n + 1) + s * m + s * 7 * n
        order = ["Qw", "qp_primal", "tempv", "dual", "w2"]
    elif Q is not None and getattr(Q, "nnz", 0) > 0:
        out["qp_primal"] = s * 12 * n + s * 2 * n1
        out["dual"] = nnz * (s + 4) + 4 * (m + 1) + s * n + s * (4 * m + 2 * m1 + m2)
        out["w2"] = nnz * (s + 4) + 4 * (n + 1) + s * m + s * 8 * n
        order = ["Qw", "qp_primal", "tempv", "dual", "w2"]
        out["Qw"] = 0
        out["tempv"] = 0
    else:
        out["primal"] = nnz * (s + 4) + 4 * (n + 1) + s * m + s * (4 * n + 2 * n1)
        out["dual"] = nnz * (s + 4) + 4 * (m + 1) + s * n + s * (3 * m + 2 * m1 + m2)
        order = ["primal", "dual"]
    return out, order


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


def cpu_baseline(model, sigma, lam_a, lam_q, iters):
    """The oracle (FP64 C restatement of the reference CPU path), 1 thread, on a
    bounded sample: `iters` HPR iterations of the same instance after setup."""
    import hprsocp
    from hprsocp import _abi
    so = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(so):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True,
                       stdout=subprocess.DEVNULL)
    olib = _abi.bind(C.CDLL(so), "oracle_")
    olib.oracle_set_threads.argtypes = [C.c_int32]
    olib.oracle_set_threads(1)
    p = hprsocp.HPRSOCP_parameters(verbose=False)
    with hprsocp.Solver(model, p, lib=olib, prefix="oracle_") as o:
        o.set_scalars(sigma=sigma, lambda_A=lam_a, lambda_Q=lam_q)
        o.update(0, False)
        t0 = time.perf_counter()
        for k in range(iters):
            o.update(k + 1, False)
        dt = time.perf_counter() - t0
    return iters / dt, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", default="c1")
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--solve-time", type=float, default=90.0, help="time cap of the time-to-1e-6 solve")
    ap.add_argument("--cpu-iters", type=int, default=5)
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32_refine"])
    ap.add_argument("--probe-copy-mb", type=int, default=0,
                    help="also run one plain device copy of this many MiB (byte-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the HPR path here is "
              "single-GPU — see the sharding design notes.", file=sys.stderr)
        sys.exit(2)
    import hprsocp
    lib = hprsocp.load_library()
    t0 = time.time()
    model, wl = make_workload(args.config, args.scale)
    m, n = model.A.shape
    log(f"workload {wl} generated in {time.time() - t0:.1f}s")
    p = hprsocp.HPRSOCP_parameters(verbose=False, stoptol=1e-6, time_limit=args.solve_time,
                                   precision=args.precision)
    t0 = time.time()
    g = hprsocp.Solver(model, p)
    log(f"setup (transfer + scaling + plans) {time.time() - t0:.2f}s")
    sc = g.estimate_lambda()
    log(f"lambda_A={sc['lambda_max_A']:.6e} lambda_Q={sc['lambda_max_Q']:.6e} sigma={sc['sigma']:.4e}")
    if args.probe_copy_mb > 0:
        pms = g.probe_copy(args.probe_copy_mb << 20)
        log(f"copy probe: {args.probe_copy_mb} MiB in {pms:.4f} ms "
            f"({2 * (args.probe_copy_mb << 20) / (pms * 1e-3) / 1e9:.0f} GB/s read+write)")
    g.time_iterations(args.warmup)
    ms, _ = g.time_iterations(args.steps)
    _, kern = g.time_iterations(max(args.steps // 4, 10), per_kernel=True)
    log(f"{ms:.4f} ms/iter -> {1000.0 / ms:.1f} iters/s; per-group ms {kern}")
    vbytes = 4 if args.precision == "fp32_refine" else 8
    bytes_by, order = algorithmic_bytes(model, vbytes)
    groups = [(name, kern[i]) for i, name in enumerate(order)]
    dom, dom_ms = max(groups, key=lambda t: t[1])
    dom_bytes = bytes_by[dom]
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    it_bytes = sum(bytes_by.values())
    traffic = MEASURED_TRAFFIC.get((args.config, args.precision, dom)) if args.scale == 1.0 else None
    g.close()

    solve = None
    if not args.no_solve:
        t0 = time.time()
        g2 = hprsocp.Solver(model, p)
        g2.estimate_lambda()
        r = g2.solve()
        g2.close()
        solve = {"status": r.status, "iter": r.iter, "time_s": r.time, "kkt": r.residuals,
                 "iter_to_1e-6": r.iter_6 if r.residuals < 1e-6 or r.iter_6 < r.iter else None,
                 "time_to_1e-6_s": r.time_6 if r.residuals < 1e-6 or r.iter_6 < r.iter else None,
                 "primal_obj": r.primal_obj, "restarts": r.restarts, "iter_fp32": r.iter_fp32}
        log(f"solve: {solve} (wall {time.time() - t0:.1f}s)")

    if args.cpu_iters > 0:
        cpu_rate, cpu_dt = cpu_baseline(model, sc["sigma"], sc["lambda_max_A"], sc["lambda_max_Q"],
                                        args.cpu_iters)
        log(f"cpu oracle: {cpu_rate:.3f} iters/s over {args.cpu_iters} iterations ({cpu_dt:.1f}s)")
    else:
        cpu_rate, cpu_dt = None, 0.0
    out = {
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
        "dtype": "f32" if args.precision == "fp32_refine" else "f64",
        "data": "synthetic (seeded feasibility construction)",
        "config": {"workload": wl, "config": args.config, "scale": args.scale, "m": m, "n": n,
                   "nnz_A": int(model.A.nnz), "precision": args.precision,
                   "algorithmic_MB_per_iter": round(it_bytes / 1e6, 2),
                   "iteration_GBps": round(it_bytes / (ms * 1e-3) / 1e9, 1)},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                     "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
                     "traffic": traffic, "algorithmic_bytes": int(dom_bytes), "ms": round(dom_ms, 5)},
        "kernel_ms": {name: round(t, 5) for name, t in groups},
        "solve_to_1e-6": solve,
        "cpu_baseline": {"value": cpu_rate and round(cpu_rate, 4), "unit": "HPR iters/s", "cores": 1,
                         "kind": "port",
                         "sample": f"{args.cpu_iters} partial HPR iterations of the same instance "
                                   f"after setup (FP64 C oracle, 1 thread, {cpu_dt:.1f}s)"},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
