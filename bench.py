#!/usr/bin/env python3
"""ALS benchmark on one MI355X: seconds per iteration at f=100, against the
plain FP32 path at equal held-out RMSE.

Workload (default): the Netflix-shaped synthetic rating matrix
m=480,189 users x n=17,770 items, 99,072,112 training ratings (power-law
degrees; rows Zipf(1.0) capped at 17,000, columns Zipf(0.8)), a 1% held-out
set, f=100, lambda=0.058.  One step = one ALS iteration = X half-sweep
(users from items over the CSR) + Theta half-sweep (items from users over
the CSC), with ratings and factors already resident in device memory.

Default path: Gram on BF16 MFMA (FP32 accumulation), FP16 conjugate
gradients with FP32 refinement and an FP64 residual, Cholesky fallback,
each row warm-started from its current factor.  The plain FP32 path (FP32
Gram + FP32 Cholesky) runs the same iterations from the same start in the
same process for the comparison; --no-plain skips it.

Prints ONE JSON line: the metric, throughput figures (s/iteration,
ratings/s, hermitian TFLOP/s), the roofline of the dominant kernel
(measured live with HIP events on the stream it runs on) and a CPU
baseline (the FP64 checker timed on this host's cores on a row sample).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(REPO, "cuda-mllib_amd"), REPO):
    if p not in sys.path:
        sys.path.insert(0, p)

METRIC = "ALS s/iteration & hermitian TFLOP/s at f=100, vs plain FP32 at equal RMSE"
PEAK_FP32_MFMA_TFLOPS = 157.3   # MI355X dense FP32 matrix peak
PEAK_BF16_MFMA_TFLOPS = 2500.0  # MI355X dense BF16 matrix peak
PEAK_HBM_GBS = 8000.0           # MI355X HBM3E


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(ds, f, lam, theta, X, target_s):
    """Times the FP64 checker's half-sweeps on row samples of both sides and
    scales to one iteration by ratings (the per-rating cost of the normal
    equations is what the sample measures)."""
    from oracle import oracle as O

    threads = O.max_threads()
    total = 0.0
    sample_desc = []
    for side, (rp, ci, cv, src) in {"x": (ds.csr_row_ptr, ds.csr_col, ds.csr_val, theta),
                                     "theta": (ds.csc_col_ptr, ds.csc_row, ds.csc_val, X)}.items():
        rows = rp.shape[0] - 1
        # Probe with a small stride sample, then size the real sample to
        # about target_s / 2 seconds per side.
        stride = max(1, rows // 200)
        for attempt in range(3):
            sel = np.arange(0, rows, stride)
            cnt = rp[sel + 1] - rp[sel]
            sub_rp = np.zeros(sel.shape[0] + 1, np.int32)
            sub_rp[1:] = np.cumsum(cnt)
            idx = np.concatenate([np.arange(rp[u], rp[u + 1]) for u in sel])
            t0 = time.perf_counter()
            O.half_sweep(sub_rp, ci[idx], cv[idx], src, lam, threads=threads)
            dt = time.perf_counter() - t0
            rate = sub_rp[-1] / dt
            if dt > 0.3 * target_s / 2 or stride == 1:
                break
            stride = max(1, int(stride * dt / (target_s / 2)))
        total += ds.nnz / rate
        sample_desc.append(f"{side}: every {stride}th row ({sel.shape[0]} rows, {int(sub_rp[-1])} ratings, {dt:.1f}s)")
    return {"value": total, "unit": "s/iteration", "cores": threads, "kind": "port",
            "sample": "FP64 normal equations + Cholesky on " + "; ".join(sample_desc)
                      + "; scaled to all 2 x 99,072,112 ratings of one iteration",
            "ratings_per_s": ds.nnz / total}


def measured_traffic(workload, precision, solver, f):
    """HBM bytes per launch of the fused kernel for this workload and path,
    from the committed rocprofv3 counter summary (FETCH_SIZE and WRITE_SIZE,
    each scaled by what the counter reads for a copy of known size); None if
    that configuration was not profiled."""
    path = os.path.join(REPO, "profiles", "traffic_per_launch.txt")
    try:
        with open(path) as fh:
            for line in fh:
                parts = line.split()
                if len(parts) == 5 and not line.startswith("#") and parts[:4] == [workload, precision, solver, str(f)]:
                    return float(parts[4])
    except OSError:
        pass
    return None


def precision_report(ds, f, lam, theta, X, nrows):
    """Error of one X half-sweep against the FP64 oracle on sampled rows:
    relative residual ||b - A x|| / ||b|| and factor error
    ||x - x*|| / ||x*|| with A, b, x* formed in FP64 from the same theta."""
    from oracle import oracle as O
    rows = np.linspace(0, ds.m - 1, nrows).astype(np.int64)
    res, err = [], []
    for u in rows:
        if ds.csr_row_ptr[u + 1] == ds.csr_row_ptr[u]:
            continue
        A, b = O.row_system(ds.csr_row_ptr, ds.csr_col, ds.csr_val, theta, f, lam, int(u))
        x = X[u].astype(np.float64)
        xs = np.linalg.solve(A, b)
        res.append(np.linalg.norm(b - A @ x) / np.linalg.norm(b))
        err.append(np.linalg.norm(x - xs) / np.linalg.norm(xs))
    q = lambda v: {"median": float(np.median(v)), "p99": float(np.percentile(v, 99)), "max": float(np.max(v))}
    return {"rows": len(res), "relative_residual": q(res), "factor_rel_error": q(err)}


def train_network(G, ds, f, lam, args, prec):
    """Warm-start network trained on pairs the plain solver produces from
    another random start on the same ratings (not timed)."""
    import cumf
    from cumf import mlp as NN
    net = NN.MLP(f, seed=1)
    hh = G.ALS(f, lam, device=0, split=args.split, precision=prec)
    hh.set_dataset(ds)
    hh.set_factors(theta=cumf.init_factors(ds.n, f, 1))
    t1 = time.perf_counter()
    net.harvest(hh, args.mlp_harvest)
    t_h = time.perf_counter() - t1
    hh.close()
    t1 = time.perf_counter()
    loss = net.train(args.mlp_epochs, batch=args.mlp_batch, lr=1e-3, seed=0)
    t_t = time.perf_counter() - t1
    learned = {"harvest_iterations": args.mlp_harvest, "pairs": net.harvested(), "harvest_s": t_h,
               "epochs": args.mlp_epochs, "batch": args.mlp_batch, "train_s": t_t, "final_loss": loss}
    log(f"warm-start network: {learned}")
    return net, learned


def run_path(G, ds, theta0, f, lam, args, prec, solver, warm, net=None, tune=True):
    """W untimed + K timed iterations of one path; returns its figures."""
    h = G.ALS(f, lam, device=0, timing=1, split=args.split, precision=prec, use_graph=int(args.graph),
              solver=solver, cg_iters=args.cg_iters, cg_rounds=args.cg_rounds, cg_tol=args.cg_tol,
              warm_start=warm)
    if net is not None:
        h.set_mlp(net)
    h.set_dataset(ds)
    h.set_factors(theta=theta0)
    tuning = None
    if tune and args.tile_table:
        h.load_tile_table(args.tile_table)
        tuning = {"tile_table": args.tile_table}
    elif tune and args.autotune:
        out = os.path.join(REPO, "gpurun_out", "tile_table_%s_f%d.txt" % (args.workload, f))
        os.makedirs(os.path.dirname(out), exist_ok=True)
        t1 = time.perf_counter()
        h.autotune(reps=2, path=out)
        tuning = {"autotune_s": time.perf_counter() - t1, "table": open(out).read().splitlines()}
        log(f"autotuned in {tuning['autotune_s']:.1f}s -> {out}")
    for _ in range(args.warmup):
        h.iterate(1)
    h.synchronize()
    h.reset_stats()
    t0 = time.perf_counter()
    h.iterate(args.steps)
    h.synchronize()
    elapsed = time.perf_counter() - t0
    st = h.stats()
    rmse = h.rmse()
    X, T = h.get_factors()
    # Per-half-sweep split, after the timed region (one more iteration).
    sweep_ms = {}
    X_next = None
    for name, fn in (("x", h.update_x), ("theta", h.update_theta)):
        h.synchronize()
        t1 = time.perf_counter()
        fn()
        h.synchronize()
        sweep_ms[name] = (time.perf_counter() - t1) * 1e3
        if name == "x" and args.precision_report:
            X_next = h.get_factors()[0]
    opt = h.opt
    h.close()
    report = precision_report(ds, f, lam, T, X_next, args.precision_report) if X_next is not None else None
    return {"s": elapsed / args.steps, "rmse": rmse, "stats": st, "sweep_ms": sweep_ms, "X": X, "T": T,
            "opt": opt, "tuning": tuning, "precision_report": report}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="netflix_480k_17770_99m")
    ap.add_argument("--f", type=int, default=100)
    ap.add_argument("--lam", type=float, default=0.058)
    ap.add_argument("--split", type=int, default=0)
    ap.add_argument("--precision", choices=["fp32", "bf16gram"], default="bf16gram")
    ap.add_argument("--solver", choices=["cholesky", "cg"], default="cg")
    ap.add_argument("--cg-iters", type=int, default=0)
    ap.add_argument("--cg-rounds", type=int, default=0)
    ap.add_argument("--cg-tol", type=float, default=0.0)
    ap.add_argument("--warm", choices=["prev", "zero", "mlp", "best"], default="prev",
                    help="CG start; mlp/best train the warm-start network first (untimed)")
    ap.add_argument("--mlp-harvest", type=int, default=3, help="plain iterations harvested for training")
    ap.add_argument("--mlp-epochs", type=int, default=4)
    ap.add_argument("--mlp-batch", type=int, default=4096)
    ap.add_argument("--graph", action="store_true", help="replay each half-sweep as a captured hipGraph")
    ap.add_argument("--autotune", action="store_true",
                    help="tune the degree-bucket launch shapes on this workload first (untimed)")
    ap.add_argument("--tile-table", default="", help="load degree-bucket launch shapes from this file")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP32 comparison run")
    ap.add_argument("--precision-report", type=int, default=0, metavar="ROWS",
                    help="per-row FP64 residual and factor error of one more X half-sweep on ROWS sampled rows")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run the plain copy kernel over BYTES bytes (HBM counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        from cumf import als as G
        ms = G.copy_kernel_ms(args.calibrate_copy, iters=5)
        print(json.dumps({"copy_bytes": args.calibrate_copy, "ms_per_launch": ms,
                          "GBps": 2 * args.calibrate_copy / ms / 1e6}), flush=True)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; the ALS sweep is single-device here "
                                   "(sharding across GPUs is described, not built)", "n_gpus": args.gpus}))
        sys.exit(2)

    import cumf
    from cumf import als as G

    t0 = time.time()
    ds = cumf.generate(cumf.preset(args.workload))
    log(f"generated {args.workload}: m={ds.m} n={ds.n} nnz={ds.nnz} test={ds.test_u.shape[0]} "
        f"in {time.time() - t0:.1f}s")
    f, lam = args.f, args.lam
    theta0 = cumf.init_factors(ds.n, f, 0)
    prec = {"fp32": G.CUMF_PREC_FP32, "bf16gram": G.CUMF_PREC_BF16_GRAM}[args.precision]
    solver = {"cholesky": G.CUMF_SOLVER_CHOLESKY, "cg": G.CUMF_SOLVER_CG}[args.solver]
    warm = {"prev": G.CUMF_WARM_PREV, "zero": G.CUMF_WARM_ZERO, "mlp": G.CUMF_WARM_MLP,
            "best": G.CUMF_WARM_BEST}[args.warm]
    if solver != G.CUMF_SOLVER_CG:
        warm = G.CUMF_WARM_PREV
    learned, net = None, None
    if warm in (G.CUMF_WARM_MLP, G.CUMF_WARM_BEST):
        net, learned = train_network(G, ds, f, lam, args, prec)

    res = run_path(G, ds, theta0, f, lam, args, prec, solver, warm, net)
    st, opt = res["stats"], res["opt"]
    log(f"timed {args.steps} iterations: {res['s']:.4f} s/iteration; rmse={res['rmse']:.5f}; stats={st}")
    is_plain = prec == G.CUMF_PREC_FP32 and solver == G.CUMF_SOLVER_CHOLESKY
    plain = None
    if not is_plain and not args.no_plain:
        plain = run_path(G, ds, theta0, f, lam, args, G.CUMF_PREC_FP32, G.CUMF_SOLVER_CHOLESKY,
                         G.CUMF_WARM_PREV, tune=False)
        log(f"plain FP32: {plain['s']:.4f} s/iteration; rmse={plain['rmse']:.5f}")

    s_per_iter = res["s"]
    # Unique symmetric Gram work: f(f+1)/2 entries, one multiply-add each, per
    # rating and half-sweep.
    gram_flops_iter = 2.0 * ds.nnz * f * (f + 1)
    # One "launch" = all als_task_kernel launches of one half-sweep (one, or
    # one per degree bucket with a tile table).
    if st["task_launches"]:  # per-kernel events
        kern_ms_total, launches, kname = st["task_ms"], st["sweeps"], "als_task_kernel (fused gram+solve)"
    else:  # graph replays: whole half-sweeps
        kern_ms_total, launches, kname = st["iter_ms"], st["iters"], "one captured half-sweep (hipGraph)"
    task_ms = kern_ms_total / max(launches, 1)
    # all kernels of the sweeps (fused Gram + solve, split-row finish, fallback)
    sweep_kernel_ms = (st["task_ms"] + st["finish_ms"] + st["fallback_ms"]) if st["task_launches"] else kern_ms_total
    traffic = measured_traffic(args.workload, args.precision, args.solver, f) if st["task_launches"] else None
    gram_flops_launch = st["gram_flops"] / max(launches, 1)
    achieved = gram_flops_launch / (task_ms * 1e-3) / 1e12
    peak = PEAK_FP32_MFMA_TFLOPS if prec == G.CUMF_PREC_FP32 else PEAK_BF16_MFMA_TFLOPS
    factor_bytes = 4 * f if prec == G.CUMF_PREC_FP32 else 2 * ((f + 7) // 8 * 8)
    gather_bytes_launch = ds.nnz * (8 + factor_bytes)  # col + rating + one factor row per rating
    if solver == G.CUMF_SOLVER_CHOLESKY:
        path = ("fp32" if prec == G.CUMF_PREC_FP32 else "bf16") + " gram (MFMA) + fp32 cholesky"
    else:
        dflt = G.default_options()  # 0 in an option means the library default
        iters, rounds = opt.cg_iters or dflt.cg_iters, opt.cg_rounds or dflt.cg_rounds
        tol = opt.cg_tol if opt.cg_tol > 0 else dflt.cg_tol
        path = (("fp32" if prec == G.CUMF_PREC_FP32 else "bf16") + " gram (MFMA, fp32 accumulate) + "
                f"fp16 cg ({iters} steps x up to {rounds} fp32 refinement rounds, fp64 acceptance residual, "
                f"tol {tol:g}, cholesky fallback, start {args.warm})")
    rec = {
        "metric": METRIC,
        "value": s_per_iter,
        "unit": "s/iteration",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": s_per_iter * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("fp32" if prec == G.CUMF_PREC_FP32 else "bf16 gram")
                 + (" / fp32 solve" if solver == G.CUMF_SOLVER_CHOLESKY
                    else " / fp16 cg + fp32 refinement + fp64 residual"),
        "data": "synthetic",
        "config": {"workload": args.workload, "m": ds.m, "n": ds.n, "nnz": ds.nnz, "f": f,
                   "lambda": lam, "path": path, "graph": bool(args.graph), "split": args.split or 16384,
                   "iterations_timed": f"{args.warmup + 1}..{args.warmup + args.steps}"},
        "ratings_per_s": ds.nnz / s_per_iter,
        "hermitian_tflops": gram_flops_iter / (sweep_kernel_ms / args.steps * 1e-3) / 1e12,
        "heldout_rmse": res["rmse"],
        "failed_rows": st["failed_rows"],
        "cg_fallback_rows_per_iter": st["cg_fallback_rows"] / args.steps,
        "kernel_ms_per_iter": {"als_task_kernel": st["task_ms"] / args.steps,
                               "als_finish_kernel": st["finish_ms"] / args.steps,
                               "als_fallback_kernel": st["fallback_ms"] / args.steps,
                               "warm_start": st["warm_ms"] / args.steps},
        "half_sweep_ms": res["sweep_ms"],
        "learned": learned,
        "batching": res["tuning"],
        "roofline": {"kernel": kname, "bound": "mfma",
                     "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": traffic,
                     "avg_launch_ms": task_ms,
                     "flops_per_launch": gram_flops_launch},
        "gather_roofline": {"bound": "hbm", "unit": "GB/s", "peak": PEAK_HBM_GBS,
                            "bytes_per_launch": gather_bytes_launch,
                            "achieved": gather_bytes_launch / (task_ms * 1e-3) / 1e9,
                            "frac": gather_bytes_launch / (task_ms * 1e-3) / 1e9 / PEAK_HBM_GBS},
    }
    if res["precision_report"] is not None:
        rec["precision_report"] = res["precision_report"]
    if plain is not None:
        rec["plain_fp32"] = {"s_per_iteration": plain["s"], "heldout_rmse": plain["rmse"],
                             "half_sweep_ms": plain["sweep_ms"]}
        if plain["precision_report"] is not None:
            rec["plain_fp32"]["precision_report"] = plain["precision_report"]
        rec["speedup_vs_plain_fp32"] = plain["s"] / s_per_iter
        rec["rmse_delta_vs_plain_fp32"] = res["rmse"] - plain["rmse"]
    if not args.no_cpu_baseline:
        log("timing CPU baseline")
        rec["cpu_baseline"] = cpu_baseline(ds, f, lam, theta0, res["X"], args.cpu_seconds)
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
