#!/usr/bin/env python3
"""ALS benchmark on one MI355X: seconds per iteration at f=100, against the
plain FP32 path at equal held-out RMSE.

Workload (default): the Netflix-shaped synthetic rating matrix
m=480,189 users x n=17,770 items, 99,072,112 training ratings (power-law
degrees; rows Zipf(1.0) capped at 17,000, columns Zipf(0.8)), a 1% held-out
set, f=100, lambda=0.058, start init_factors(17770, 100, seed 0).  One step
= one ALS iteration = X half-sweep (users from items over the CSR) + Theta
half-sweep (items from users over the CSC), with ratings and factors
resident in device memory.

Default path (the combined form): Gram on BF16 MFMA with FP32
accumulation; per degree bucket a launch shape and a solver (FP32 Cholesky,
or FP16 conjugate gradients inside FP32 refinement with an FP64 residual
and Cholesky fallback) chosen by the library's autotuner on this workload
(untimed); every row solved to a relative residual of 1e-5 against the
Gram the GPU formed; each half-sweep replayed from a captured hipGraph;
CG started from the row's current factor.

Per path, in order:
  1. training: 10 iterations from the seeded start, timed as a whole, and
     the held-out RMSE after them against the FP64 oracle's 10-iteration
     RMSE where a fixture holds it (tests/golden/netflix_f100_fp64.npz);
  2. W untimed warm-up iterations, then K timed iterations (value);
  3. per-kernel times of the same launches without graph replay (one
     more iteration with HIP events) for the roofline, then one X and one
     Theta half-sweep timed on their own.
The plain FP32 path (FP32 Gram + FP32 Cholesky, the drop-in's arithmetic)
runs the same steps from the same start for the comparison; the drop-in
entry point doALSWithCSR (host factors in and out, PCIe included) is timed
on one X and one Theta half-sweep.

Prints ONE JSON line.
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import tempfile
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
FIXTURES = {("netflix_480k_17770_99m", 100, 0.058): os.path.join(REPO, "tests", "golden", "netflix_f100_fp64.npz")}
TRAIN_ITERS = 10


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return platform.processor() or "unknown"


def cpu_baseline(ds, f, lam, theta, X, target_s):
    """The FP64 normal-equation path (the oracle's C code, OpenMP over every
    host core) timed on this host: one X half-sweep and one Theta
    half-sweep, each on all rows if that fits in about target_s / 2 seconds,
    else on an evenly strided row sample scaled by ratings."""
    from oracle import oracle as O

    threads = O.max_threads()
    total = 0.0
    sample_desc = []
    for side, (rp, ci, cv, src) in {"x": (ds.csr_row_ptr, ds.csr_col, ds.csr_val, theta),
                                     "theta": (ds.csc_col_ptr, ds.csc_row, ds.csc_val, X)}.items():
        rows = rp.shape[0] - 1
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
        what = "all rows" if stride == 1 else f"every {stride}th row"
        sample_desc.append(f"{side}: {what} ({sel.shape[0]} rows, {int(sub_rp[-1])} ratings, {dt:.1f}s)")
    return {"value": total, "unit": "s/iteration", "cores": threads, "kind": "port", "cpu": cpu_model(),
            "sample": "FP64 normal equations + Cholesky (the oracle's C restatement) on " + "; ".join(sample_desc)
                      + "; scaled by ratings to one iteration of 2 x %d ratings" % ds.nnz,
            "ratings_per_s": ds.nnz / total}


# What the committed counter passes show limits the dominant kernel, per
# profiled configuration (profiles/als_task_kernel_netflix480k_f100_counters.md).
LIMITER = {
    ("netflix_480k_17770_99m", "bf16-cg-tuned", 100):
        "X half-sweep (dual form up to 95 ratings, Cholesky above): latency of gathers and of the "
        "per-row solves (35% of wave time waiting on memory, 37% waiting to issue, matrix pipe 32% busy); "
        "Theta half-sweep: BF16 gathers beyond L2 (48% of wave time waiting on memory)",
}


def measured_traffic(workload, path, f):
    """HBM bytes per launch of the fused kernel for this workload and path,
    from the committed rocprofv3 counter summary (FETCH_SIZE and WRITE_SIZE,
    each scaled by what the counter reads for a copy of known size); None if
    that configuration was not profiled."""
    fn = os.path.join(REPO, "profiles", "traffic_per_launch.txt")
    try:
        with open(fn) as fh:
            for line in fh:
                parts = line.split()
                if len(parts) == 4 and not line.startswith("#") and parts[:3] == [workload, path, str(f)]:
                    return float(parts[3])
    except OSError:
        pass
    return None


def _bf16_round(a):
    """FP32 -> BF16 (nearest even) -> FP32: the factor copy a BF16 Gram reads."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32).view(np.float32)


def precision_report(ds, f, lam, theta, X, nrows, bf16):
    """Error of one X half-sweep on sampled rows (the heaviest and lightest
    eighths and a seeded uniform draw), against the FP64 normal equations
    formed here with numpy: A = Y^T Y + lambda n_u I, b = Y^T r from the same
    factors -- as the GPU read them (BF16-rounded for a BF16 Gram) and as
    FP32.  Relative residual ||b - A x|| / ||b|| and factor error
    ||x - x*|| / ||x*||, x* = A^-1 b in FP64."""
    rp = ds.csr_row_ptr
    deg = np.diff(rp)
    live = np.flatnonzero(deg > 0)
    order = live[np.argsort(-deg[live], kind="stable")]
    pick = list(order[:nrows // 8]) + list(order[-(nrows // 8):])
    rest = np.setdiff1d(live, pick)
    pick += list(np.random.default_rng(11).choice(rest, nrows - len(pick), replace=False))
    rows = np.sort(np.array(pick, np.int64))

    def errors(src):
        res, err = [], []
        for u in rows:
            s, e = rp[u], rp[u + 1]
            Y = src[ds.csr_col[s:e]].astype(np.float64)
            r = ds.csr_val[s:e].astype(np.float64)
            A = Y.T @ Y + np.float64(np.float32(lam) * np.float32(e - s)) * np.eye(f)
            b = Y.T @ r
            x = X[u].astype(np.float64)
            xs = np.linalg.solve(A, b)
            res.append(np.linalg.norm(b - A @ x) / np.linalg.norm(b))
            err.append(np.linalg.norm(x - xs) / np.linalg.norm(xs))
        return res, err

    q = lambda v: {"median": float(np.median(v)), "p99": float(np.percentile(v, 99)), "max": float(np.max(v))}
    out = {"rows": int(rows.shape[0]), "sample": "heaviest and lightest eighths + seeded uniform draw"}
    if bf16:
        res, err = errors(_bf16_round(theta))
        out["vs_system_on_bf16_factors"] = {"relative_residual": q(res), "factor_rel_error": q(err)}
    res, err = errors(theta)
    out["vs_system_on_fp32_factors"] = {"relative_residual": q(res), "factor_rel_error": q(err)}
    return out


def train_network(G, ds, f, lam, args, prec):
    """Warm-start network trained on pairs the plain solver produces from
    another random start on the same ratings (not timed)."""
    import cumf
    from cumf import mlp as NN
    net = NN.MLP(f, seed=1)
    hh = G.ALS(f, lam, device=0, precision=prec)
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


def learned_report(G, ds, theta0, f, lam, args, prec, fp64_rmse):
    """The learned warm-start form at its configured setting: CG capped at 3
    steps in one round, Cholesky whenever the relative residual stays above
    1e-5.  A 10-iteration training from the seeded start with the network's
    prediction as the CG start, and the same with the previous factor as the
    start (the form without the network); untimed network training first."""
    net, info = train_network(G, ds, f, lam, args, prec)
    out = {"setting": "cg_iters 3, cg_rounds 1, cg_tol 1e-5, Cholesky fallback", "network": info}
    for name, warm in (("with_network", G.CUMF_WARM_MLP), ("without_network_prev_start", G.CUMF_WARM_PREV)):
        h = G.ALS(f, lam, device=0, precision=prec, solver=G.CUMF_SOLVER_CG, cg_iters=3, cg_rounds=1,
                  cg_tol=1e-5, warm_start=warm)
        if warm == G.CUMF_WARM_MLP:
            h.set_mlp(net)
        h.set_dataset(ds)
        h.set_factors(theta=theta0)
        h.synchronize()
        t0 = time.perf_counter()
        h.iterate(TRAIN_ITERS)
        h.synchronize()
        t = time.perf_counter() - t0
        rm = h.rmse()
        fb = h.stats()["cg_fallback_rows"]
        h.close()
        out[name] = {"train_s": t, "heldout_rmse": rm,
                     "rmse_delta_vs_fp64_oracle": (rm - fp64_rmse) if fp64_rmse else None,
                     "cg_fallback_rows": fb, "fallback_rate": fb / (TRAIN_ITERS * (ds.m + ds.n))}
    net.close()
    log(f"learned form: {out}")
    return out


def run_path(G, ds, theta0, f, lam, args, name, opts, tune, net=None, report_rows=0):
    """Training, steady-state timing and per-kernel split of one path."""
    h = G.ALS(f, lam, device=0, **opts)
    if net is not None:
        h.set_mlp(net)
    h.set_dataset(ds)
    h.set_factors(theta=theta0)
    tuning = None
    if tune == "autotune":
        out = os.path.join(args.out_dir, "tile_table_%s_f%d.txt" % (args.workload, f))
        os.makedirs(os.path.dirname(out), exist_ok=True)
        t1 = time.perf_counter()
        h.autotune(reps=2, path=out)
        tuning = {"autotune_s": time.perf_counter() - t1, "table": [l for l in open(out).read().splitlines()
                                                                    if not l.startswith("#")]}
        log(f"{name}: autotuned in {tuning['autotune_s']:.1f}s -> {out}")
    elif tune:
        h.load_tile_table(tune)
        tuning = {"tile_table": tune}
    # 1. training from the seeded start
    h.synchronize()
    t0 = time.perf_counter()
    h.iterate(TRAIN_ITERS)
    h.synchronize()
    train_s = time.perf_counter() - t0
    train_rmse = h.rmse()
    train_fallback = h.stats()["cg_fallback_rows"]
    # 2. steady state: W warm-up + K timed iterations; with --retune the
    #    tile table is tuned again on the trained factors first (a CG bucket
    #    from a converged start needs fewer rounds than from the cold one).
    if tune == "autotune" and args.retune:
        out = os.path.join(args.out_dir, "tile_table_%s_f%d.txt" % (args.workload, f))
        t1 = time.perf_counter()
        h.autotune(reps=2, path=out)
        tuning["retune_s"] = time.perf_counter() - t1
        tuning["table"] = [l for l in open(out).read().splitlines() if not l.startswith("#")]
        log(f"{name}: tuned again after training in {tuning['retune_s']:.1f}s")
    for _ in range(args.warmup):
        h.iterate(1)
    h.synchronize()
    h.reset_stats()
    t0 = time.perf_counter()
    h.iterate(args.steps)
    h.synchronize()
    elapsed = time.perf_counter() - t0
    st = h.stats()
    X, T = h.get_factors()
    h.close()
    # 3. per-kernel times without graph replay, and the half-sweep split
    eo = dict(opts, use_graph=0, timing=1)
    he = G.ALS(f, lam, device=0, **eo)
    if net is not None:
        he.set_mlp(net)
    he.set_dataset(ds)
    if tune == "autotune":
        he.load_tile_table(os.path.join(args.out_dir, "tile_table_%s_f%d.txt" % (args.workload, f)))
    elif tune:
        he.load_tile_table(tune)
    he.set_factors(X=X, theta=T)
    he.iterate(1)
    he.synchronize()
    he.reset_stats()
    he.iterate(1)
    he.synchronize()
    kst = he.stats()
    sweep_ms, report = {}, None
    for side, fn in (("x", he.update_x), ("theta", he.update_theta)):
        he.synchronize()
        t1 = time.perf_counter()
        fn()
        he.synchronize()
        sweep_ms[side] = (time.perf_counter() - t1) * 1e3
        if side == "x" and report_rows:
            Xn, Tn = he.get_factors()
            report = precision_report(ds, f, lam, Tn, Xn, report_rows, opts.get("precision", 0) != 0)
    opt = he.opt
    he.close()
    res = {"s": elapsed / args.steps, "stats": st, "kernel_stats": kst, "sweep_ms": sweep_ms, "X": X, "T": T,
           "opt": opt, "tuning": tuning, "precision_report": report, "train_s": train_s,
           "train_rmse": train_rmse, "train_fallback": train_fallback, "rmse": None}
    log(f"{name}: train {TRAIN_ITERS} it {train_s:.3f}s rmse {train_rmse:.6f}; steady {res['s'] * 1e3:.3f} ms/it; "
        f"half-sweeps {sweep_ms}")
    return res


def time_dropin(G, ds, f, lam, theta, X):
    """doALSWithCSR (host factors in and out) on one X and one Theta
    half-sweep, the second call of each timed (the first fills the per-device
    plan cache); the caller's CSR upload into mallocBest memory is timed on
    its own."""
    out = {}
    for side, (rp, ci, cv, src, rows, other) in {
            "x": (ds.csr_row_ptr, ds.csr_col, ds.csr_val, theta, ds.m, ds.n),
            "theta": (ds.csc_col_ptr, ds.csc_row, ds.csc_val, X, ds.n, ds.m)}.items():
        t_call = t_up = 0.0
        for rep in range(2):
            t0 = time.perf_counter()
            bufs = [G.DeviceBuffer.from_array(a) for a in (rp, ci, cv)]
            t1 = time.perf_counter()
            Y = np.zeros((rows, f), np.float32)
            G.lib().doALSWithCSR(None, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, G._ptr(np.ascontiguousarray(src)),
                                 G._ptr(Y), rows, other, f, int(ci.shape[0]), lam, 1)
            t2 = time.perf_counter()
            for b in bufs:
                b.ptr = G.P()
            err = G.last_error()
            if err:
                raise G.CumfError(f"doALSWithCSR: {err}")
            t_up, t_call = t1 - t0, t2 - t1
        out[side] = {"call_ms": t_call * 1e3, "csr_upload_ms": t_up * 1e3}
    out["ms_per_iteration"] = out["x"]["call_ms"] + out["theta"]["call_ms"]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="netflix_480k_17770_99m")
    ap.add_argument("--f", type=int, default=100)
    ap.add_argument("--lam", type=float, default=0.058)
    ap.add_argument("--split", type=int, default=0)
    ap.add_argument("--precision", choices=["fp32", "bf16gram"], default="bf16gram")
    ap.add_argument("--solver", choices=["cholesky", "cg"], default="cg",
                    help="the handle's solver; with cg the autotuner may pick Cholesky per degree bucket")
    ap.add_argument("--cg-iters", type=int, default=0)
    ap.add_argument("--cg-rounds", type=int, default=0)
    ap.add_argument("--cg-tol", type=float, default=0.0, help="0: the library default (1e-5)")
    ap.add_argument("--warm", choices=["prev", "zero", "mlp", "best"], default="prev",
                    help="CG start; mlp/best train the warm-start network first (untimed)")
    ap.add_argument("--mlp-harvest", type=int, default=5, help="plain iterations harvested for training")
    ap.add_argument("--mlp-epochs", type=int, default=2)
    ap.add_argument("--mlp-batch", type=int, default=4096)
    ap.add_argument("--no-graph", action="store_true", help="launch each half-sweep eagerly")
    ap.add_argument("--no-tune", action="store_true", help="one launch per kernel, no tile table")
    ap.add_argument("--tile-table", default="", help="load degree-bucket launch shapes from this file")
    ap.add_argument("--out-dir", default=os.path.join(tempfile.gettempdir(), "cumf_bench"),
                    help="folder the autotuned tile table is written to")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP32 comparison run")
    ap.add_argument("--no-dropin", action="store_true", help="skip timing doALSWithCSR")
    ap.add_argument("--retune", action="store_true",
                    help="tune the tile table again on the trained factors before the timed iterations")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned warm-start form at its configured setting (3 CG steps, tol 1e-5)")
    ap.add_argument("--precision-report", type=int, default=256, metavar="ROWS",
                    help="FP64 residual and factor error of one X half-sweep on ROWS sampled rows (0: off)")
    ap.add_argument("--cpu-seconds", type=float, default=24.0)
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
        print(json.dumps({"error": "the ALS sweep runs on one GPU; sharding across GPUs is not implemented, "
                                   "so only --gpus 1 is measured", "n_gpus": args.gpus}))
        sys.exit(2)

    import cumf
    from cumf import als as G

    t0 = time.time()
    ds = cumf.generate(cumf.preset(args.workload))
    log(f"generated {args.workload}: m={ds.m} n={ds.n} nnz={ds.nnz} test={ds.test_u.shape[0]} "
        f"in {time.time() - t0:.1f}s")
    f, lam = args.f, args.lam
    theta0 = cumf.init_factors(ds.n, f, 0)
    fx_path = FIXTURES.get((args.workload, f, round(lam, 6)))
    fp64_rmse = float(np.load(fx_path)["rmse"][TRAIN_ITERS - 1]) if fx_path and os.path.exists(fx_path) else None
    prec = {"fp32": G.CUMF_PREC_FP32, "bf16gram": G.CUMF_PREC_BF16_GRAM}[args.precision]
    solver = {"cholesky": G.CUMF_SOLVER_CHOLESKY, "cg": G.CUMF_SOLVER_CG}[args.solver]
    warm = {"prev": G.CUMF_WARM_PREV, "zero": G.CUMF_WARM_ZERO, "mlp": G.CUMF_WARM_MLP,
            "best": G.CUMF_WARM_BEST}[args.warm]
    if solver != G.CUMF_SOLVER_CG:
        warm = G.CUMF_WARM_PREV
    learned, net = None, None
    if warm in (G.CUMF_WARM_MLP, G.CUMF_WARM_BEST):
        net, learned = train_network(G, ds, f, lam, args, prec)
    opts = dict(precision=prec, solver=solver, split=args.split, cg_iters=args.cg_iters, cg_rounds=args.cg_rounds,
                cg_tol=args.cg_tol, warm_start=warm, use_graph=0 if args.no_graph else 1, timing=0)
    tune = None if args.no_tune else (args.tile_table or "autotune")
    res = run_path(G, ds, theta0, f, lam, args, "path", opts, tune, net, args.precision_report)
    st, kst, opt = res["stats"], res["kernel_stats"], res["opt"]
    if learned is not None:
        learned["cg_fallback_rows_training"] = res["train_fallback"]
    elif not args.no_learned and f <= 140:
        learned = learned_report(G, ds, theta0, f, lam, args, prec, fp64_rmse)
    is_plain = prec == G.CUMF_PREC_FP32 and solver == G.CUMF_SOLVER_CHOLESKY
    plain = None
    if not is_plain and not args.no_plain:
        popts = dict(precision=G.CUMF_PREC_FP32, solver=G.CUMF_SOLVER_CHOLESKY, split=args.split,
                     use_graph=0 if args.no_graph else 1, timing=0)
        plain = run_path(G, ds, theta0, f, lam, args, "plain fp32", popts, None, None, args.precision_report)

    s_per_iter = res["s"]
    # Unique symmetric Gram work: f(f+1)/2 entries, one multiply-add each, per
    # rating and half-sweep.
    gram_flops_iter = 2.0 * ds.nnz * f * (f + 1)
    # The dominant kernel: als_task_kernel, all its launches of one half-sweep
    # (one, or one per degree bucket) = one "launch"; per-kernel HIP events of
    # the eager iteration after the timed region (graph replays carry none).
    launches = max(kst["sweeps"], 1)
    task_ms = kst["task_ms"] / launches
    sweep_kernel_ms = (kst["task_ms"] + kst["finish_ms"] + kst["fallback_ms"]) / launches * 2
    dflt = G.default_options()  # 0 in an option means the library default
    iters, rounds = opt.cg_iters or dflt.cg_iters, opt.cg_rounds or dflt.cg_rounds
    tol = opt.cg_tol if opt.cg_tol > 0 else dflt.cg_tol
    pname = "bf16" if prec != G.CUMF_PREC_FP32 else "fp32"
    traffic_key = f"{pname}-{args.solver}" + ("" if args.no_tune else "-tuned")
    traffic = measured_traffic(args.workload, traffic_key, f)
    gram_flops_launch = kst["gram_flops"] / launches
    achieved = gram_flops_launch / (task_ms * 1e-3) / 1e12
    peak = PEAK_FP32_MFMA_TFLOPS if prec == G.CUMF_PREC_FP32 else PEAK_BF16_MFMA_TFLOPS
    factor_bytes = 4 * f if prec == G.CUMF_PREC_FP32 else 2 * ((f + 7) // 8 * 8)
    gather_bytes_launch = ds.nnz * (8 + factor_bytes)  # col + rating + one factor row per rating
    if solver == G.CUMF_SOLVER_CHOLESKY:
        path = f"{pname} gram (MFMA, fp32 accumulate) + fp32 cholesky (MFMA panel and trailing updates)"
    else:
        path = (f"{pname} gram (MFMA, fp32 accumulate) + per degree bucket fp32 cholesky or fp16 cg "
                f"({iters} steps x up to {rounds} fp32 refinement rounds, fp64 residual, start {args.warm}) "
                "as tuned; every row to relative residual <= %g against the GPU's Gram" % tol)
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
        "dtype": f"{'bf16' if prec else 'fp32'} gram / fp32 accumulate"
                 + (" / fp32 cholesky" if solver == G.CUMF_SOLVER_CHOLESKY
                    else " / fp32 cholesky or fp16 cg + fp32 refinement + fp64 residual"),
        "data": "synthetic (seeded generator: power-law degrees, rank-20 latent ratings 1-5); "
                "factors from init_factors seed 0",
        "config": {"workload": args.workload, "m": ds.m, "n": ds.n, "nnz": ds.nnz, "f": f,
                   "lambda": lam, "path": path, "graph": not args.no_graph, "split": opt.split or dflt.split,
                   "tol": tol if solver == G.CUMF_SOLVER_CG else None,
                   "iterations_timed": f"{TRAIN_ITERS + args.warmup + 1}..{TRAIN_ITERS + args.warmup + args.steps}"},
        "ratings_per_s": ds.nnz / s_per_iter,
        "hermitian_tflops": gram_flops_iter / (sweep_kernel_ms * 1e-3) / 1e12,
        "training": {"iterations": TRAIN_ITERS, "s": res["train_s"], "heldout_rmse": res["train_rmse"],
                     "fp64_oracle_rmse": fp64_rmse,
                     "rmse_delta_vs_fp64_oracle": (res["train_rmse"] - fp64_rmse) if fp64_rmse else None,
                     "cg_fallback_rows": res["train_fallback"]},
        "failed_rows": st["failed_rows"],
        "cg_fallback_rows_per_iter": st["cg_fallback_rows"] / args.steps,
        "kernel_ms_per_iter": {"als_task_kernel": kst["task_ms"],
                               "als_finish_kernel": kst["finish_ms"],
                               "als_fallback_kernel": kst["fallback_ms"],
                               "warm_start": kst["warm_ms"]},
        "half_sweep_ms": res["sweep_ms"],
        "learned": learned,
        "batching": res["tuning"],
        "roofline": {"kernel": ("als_task_kernel (fused gram+solve)" if f <= 140 else
                                "als_large_kernel + als_large_dual_kernel (work-group per row; dual form "
                                "for rows of at most 95 ratings)") + ", per half-sweep", "bound": "mfma",
                     "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": traffic,
                     "avg_launch_ms": task_ms,
                     "flops_per_launch": gram_flops_launch,
                     "limiter": LIMITER.get((args.workload, traffic_key, f))},
        "gather_roofline": {"bound": "hbm", "unit": "GB/s", "peak": PEAK_HBM_GBS,
                            "bytes_per_launch": gather_bytes_launch,
                            "achieved": gather_bytes_launch / (task_ms * 1e-3) / 1e9,
                            "frac": gather_bytes_launch / (task_ms * 1e-3) / 1e9 / PEAK_HBM_GBS},
    }
    if res["precision_report"] is not None:
        rec["precision_report"] = res["precision_report"]
    if plain is not None:
        rec["plain_fp32"] = {"s_per_iteration": plain["s"], "half_sweep_ms": plain["sweep_ms"],
                             "training": {"iterations": TRAIN_ITERS, "s": plain["train_s"],
                                          "heldout_rmse": plain["train_rmse"],
                                          "rmse_delta_vs_fp64_oracle":
                                              (plain["train_rmse"] - fp64_rmse) if fp64_rmse else None},
                             "als_task_kernel_ms_per_iter": plain["kernel_stats"]["task_ms"]}
        if plain["precision_report"] is not None:
            rec["plain_fp32"]["precision_report"] = plain["precision_report"]
        rec["speedup_vs_plain_fp32"] = plain["s"] / s_per_iter
        rec["training_speedup_vs_plain_fp32"] = plain["train_s"] / res["train_s"]
        rec["rmse_delta_vs_plain_fp32"] = res["train_rmse"] - plain["train_rmse"]
    if not args.no_dropin:
        log("timing doALSWithCSR")
        rec["dropin_doALSWithCSR"] = time_dropin(G, ds, f, lam, theta0, res["X"])
        rec["dropin_doALSWithCSR"]["note"] = ("legacy C ABI, FP32 Cholesky, host factors in and out "
                                              "(PCIe included); not the value")
    if not args.no_cpu_baseline:
        log("timing CPU baseline")
        rec["cpu_baseline"] = cpu_baseline(ds, f, lam, theta0, res["X"], args.cpu_seconds)
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
