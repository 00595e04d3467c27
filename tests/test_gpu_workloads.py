"""Parity on the benchmark workloads at full size, through the C ABI.

Netflix-shaped workload (generator preset netflix_480k_17770_99m: 480,189 x
17,770, 99,072,112 training ratings, 990,721 held out), f=100,
lambda=0.058, start init_factors(17770, 100, seed 0).  Each path runs two
iterations, one more X and one more Theta half-sweep, and seven more
iterations (ten in all):
  * the factors after two iterations are compared with the FP64 oracle's
    trajectory on a fixed row sample (tests/golden/netflix_f100_fp64.npz,
    written by tests/golden/make_workload_fixtures.py);
  * the third X and Theta half-sweeps are checked row by row, on rows
    sampled across the degree range, against the FP64 oracle solving from
    the same factors: relative residual ||b - A x|| / ||b|| and factor error;
  * the held-out RMSE after ten iterations is compared with the FP64
    oracle's (fixture), within 1e-3.
Plain FP32 (FP32 Gram + FP32 Cholesky): residual <= 1e-5 and factor error
<= 1e-3 against the FP64 system on the same FP32 factors.  Precision form
(BF16 Gram): residual <= 1e-5 (Cholesky) or <= 2e-5 (FP16 CG at tol 1e-5)
against the FP64 system on the BF16-rounded factors the Gram reads; the
error against the system on the FP32 factors (what BF16 rounding costs) is
bounded and printed.

4M x 500k x 1.0e9 workload (preset zipf_4m_500k_1b), f=128, lambda=0.05,
batching form (per-bucket persistent launches from a tile table, each
half-sweep replayed as a captured graph, BF16 Gram, FP16 CG at tol 1e-5
with Cholesky fallback): two iterations, then
one X and one Theta half-sweep checked on sampled rows against the oracle.

Learned warm-start form at the Netflix shape: network trained on the X
half-sweeps of five plain iterations, CG capped at three steps from its
prediction, Cholesky when the relative residual stays above 1e-5; the
fallback rate is reported, every solve meets 1e-5 and the RMSE matches the
FP64 oracle's.
"""
import os
import time

import numpy as np
import pytest

import cumf
from cumf import als as G
from cumf import mlp as NN
from tests._checks import bf16_round, dataset_crc, degree_sample, rel_rows, row_errors

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "netflix_f100_fp64.npz")
F, LAM = 100, 0.058


@pytest.fixture(scope="module")
def netflix():
    t0 = time.time()
    ds = cumf.generate(cumf.preset("netflix_480k_17770_99m"))
    fx = np.load(FIXTURE)
    assert np.array_equal(dataset_crc(ds), fx["crc"]), \
        "the generator's Netflix-shaped ratings differ from those the FP64 fixture was made from"
    print(f"netflix: generated in {time.time() - t0:.1f}s")
    return ds, fx


def _trajectory(ds, table=None, **opt):
    """2 iterations, X half-sweep, Theta half-sweep, 7 iterations."""
    T0 = cumf.init_factors(ds.n, F, 0)
    with G.ALS(F, LAM, **opt) as h:
        h.set_dataset(ds)
        if table is not None:
            h.load_tile_table(str(table))
        h.set_factors(theta=T0)
        h.iterate(2)
        X2, T2 = h.get_factors()
        h.update_x()
        X3, _ = h.get_factors()
        h.update_theta()
        _, T3 = h.get_factors()
        h.iterate(7)
        rmse = h.rmse()
        st = h.stats()
    return X2, T2, X3, T3, rmse, st


def _q(v):
    return f"median {np.median(v):.2e} p99 {np.percentile(v, 99):.2e} max {np.max(v):.2e}"


def _third_sweep(ds, X2, T2, X3, T3, bf16):
    """Row checks of the third half-sweeps; returns (residual, error) arrays of
    the X and Theta samples against the systems on the factors the Gram
    read, and against those on the FP32 factors."""
    users = degree_sample(ds.csr_row_ptr, 256, 3)
    items = degree_sample(ds.csc_col_ptr, 48, 4)
    src_t = bf16_round(T2) if bf16 else T2
    src_x = bf16_round(X3) if bf16 else X3
    out = {"x": row_errors(ds.csr_row_ptr, ds.csr_col, ds.csr_val, src_t, X3, F, LAM, users),
           "theta": row_errors(ds.csc_col_ptr, ds.csc_row, ds.csc_val, src_x, T3, F, LAM, items)}
    if bf16:
        out["x32"] = row_errors(ds.csr_row_ptr, ds.csr_col, ds.csr_val, T2, X3, F, LAM, users)
        out["theta32"] = row_errors(ds.csc_col_ptr, ds.csc_row, ds.csc_val, X3, T3, F, LAM, items)
    for k, (res, err) in out.items():
        print(f"  {k}: residual {_q(res)}; factor error {_q(err)}")
    return out


def test_netflix_plain_fp32_vs_fp64_oracle(netflix):
    ds, fx = netflix
    X2, T2, X3, T3, rmse, st = _trajectory(ds, precision=G.CUMF_PREC_FP32, solver=G.CUMF_SOLVER_CHOLESKY)
    assert st["failed_rows"] == 0
    dx = rel_rows(X2[fx["users"]], fx["X2"])
    dt = rel_rows(T2[fx["items"]], fx["T2"])
    print(f"plain: after 2 iterations vs FP64 trajectory: X {_q(dx)}; Theta {_q(dt)}")
    out = _third_sweep(ds, X2, T2, X3, T3, False)
    for k in ("x", "theta"):
        res, err = out[k]
        assert res.max() <= 1e-5, (k, res.max())
        assert err.max() <= 1e-3, (k, err.max())
    assert np.median(dx) <= 1e-3 and np.median(dt) <= 1e-3, (np.median(dx), np.median(dt))
    print(f"plain: held-out RMSE after 10 iterations {rmse:.6f}, FP64 oracle {fx['rmse'][9]:.6f}")
    assert abs(rmse - fx["rmse"][9]) <= 1e-3


@pytest.mark.parametrize("solver", ["cholesky", "cg", "dual"])
def test_netflix_precision_form_vs_fp64_oracle(netflix, solver, tmp_path):
    """solver "dual": Cholesky handle whose rows of at most 95 ratings (degree
    buckets 0-2, 314,655 of the 480,189 users) take the dual form."""
    ds, fx = netflix
    opt = dict(precision=G.CUMF_PREC_BF16_GRAM)
    table = None
    if solver == "cg":
        opt.update(solver=G.CUMF_SOLVER_CG, cg_iters=6, cg_tol=1e-5)
    else:
        opt.update(solver=G.CUMF_SOLVER_CHOLESKY)
    if solver == "dual":
        table = tmp_path / "dual.txt"
        table.write_text("".join(f"bucket {s} {b} 0 0 variant 2 persistent 0 solver {2 if b < 3 else -1}\n"
                                 for s in (0, 1) for b in range(8)))
    X2, T2, X3, T3, rmse, st = _trajectory(ds, table=table, **opt)
    assert st["failed_rows"] == 0
    dx = rel_rows(X2[fx["users"]], fx["X2"])
    dt = rel_rows(T2[fx["items"]], fx["T2"])
    print(f"bf16 gram + {solver}: after 2 iterations vs FP64 trajectory: X {_q(dx)}; Theta {_q(dt)}; "
          f"CG fallback rows {st['cg_fallback_rows']}")
    out = _third_sweep(ds, X2, T2, X3, T3, True)
    bar = 2e-5 if solver == "cg" else 1e-5
    for k in ("x", "theta"):
        res, err = out[k]
        assert res.max() <= bar, (k, res.max())
        # Against the system on the FP32 factors: BF16 rounding of the
        # gathered factor rows (relative 2^-9 per entry) is what is left.
        res32, err32 = out[k + "32"]
        assert res32.max() <= 2e-3 and err32.max() <= 5e-2, (k, res32.max(), err32.max())
    print(f"bf16 gram + {solver}: held-out RMSE after 10 iterations {rmse:.6f}, FP64 oracle {fx['rmse'][9]:.6f}")
    assert abs(rmse - fx["rmse"][9]) <= 1e-3


def test_netflix_learned_form_at_its_configured_setting(netflix):
    """Network trained on 5 plain iterations' X half-sweeps (2.4M pairs), CG
    capped at 3 steps from its prediction, Cholesky above 1e-5."""
    ds, fx = netflix
    T0 = cumf.init_factors(ds.n, F, 0)
    net = NN.MLP(F, seed=1)
    with G.ALS(F, LAM) as hh:
        hh.set_dataset(ds)
        hh.set_factors(theta=cumf.init_factors(ds.n, F, 1))
        net.harvest(hh, 5)
    assert net.harvested() == 5 * ds.m
    loss = net.train(2, batch=4096, lr=1e-3, seed=0)
    with G.ALS(F, LAM, solver=G.CUMF_SOLVER_CG, cg_iters=3, cg_rounds=1, cg_tol=1e-5,
               warm_start=G.CUMF_WARM_MLP) as h:
        h.set_mlp(net)
        h.set_dataset(ds)
        h.set_factors(theta=T0)
        h.iterate(2)
        X2, T2 = h.get_factors()
        h.reset_stats()
        h.update_x()
        X3, _ = h.get_factors()
        st = h.stats()
        h.update_theta()
        _, T3 = h.get_factors()
        h.iterate(7)
        rmse = h.rmse()
    net.close()
    rate = st["cg_fallback_rows"] / ds.m
    print(f"learned: training loss {loss:.4g}; X half-sweep fallback rate {rate:.4f} "
          f"({st['cg_fallback_rows']} of {ds.m} rows); RMSE {rmse:.6f} vs FP64 {fx['rmse'][9]:.6f}")
    assert 0.0 <= rate <= 1.0
    users = degree_sample(ds.csr_row_ptr, 256, 5)
    res, err = row_errors(ds.csr_row_ptr, ds.csr_col, ds.csr_val, T2, X3, F, LAM, users)
    print(f"  x: residual {_q(res)}; factor error {_q(err)}")
    # Accepted rows met 1e-5 against the FP32 Gram, the rest were solved by
    # the FP32 Cholesky: the FP64 system sees either within 2e-5.
    assert res.max() <= 2e-5, res.max()
    assert abs(rmse - fx["rmse"][9]) <= 1e-3


TILE_TABLE_F128 = """# per-bucket launch shapes for the 1e9-rating workload at f=128
side 0 split 16384 finish 1 fallback 1
side 1 split 16384 finish 1 fallback 1
""" + "".join(f"bucket {s} {b} 0 0 variant 1 persistent {2048 if b < 4 else 0}\n"
              for s in (0, 1) for b in range(8))


@pytest.mark.timeout(900)
def test_zipf_1e9_batching_form_vs_oracle(tmp_path):
    t0 = time.time()
    ds = cumf.generate(cumf.preset("zipf_4m_500k_1b"))
    print(f"zipf_4m_500k_1b: m={ds.m} n={ds.n} nnz={ds.nnz} generated in {time.time() - t0:.1f}s", flush=True)
    f, lam = 128, 0.05
    table = tmp_path / "tiles.txt"
    table.write_text(TILE_TABLE_F128)
    T0 = cumf.init_factors(ds.n, f, 0)
    with G.ALS(f, lam, precision=G.CUMF_PREC_BF16_GRAM, solver=G.CUMF_SOLVER_CG, cg_tol=1e-5, use_graph=1) as h:
        h.set_dataset(ds)
        h.load_tile_table(str(table))
        h.set_factors(theta=T0)
        h.iterate(2)
        _, T2 = h.get_factors()
        h.update_x()
        X3, _ = h.get_factors()
        h.update_theta()
        _, T3 = h.get_factors()
        st = h.stats()
        rmse = h.rmse()
    print(f"zipf_4m_500k_1b: 2 iterations + 2 half-sweeps done at {time.time() - t0:.1f}s; RMSE {rmse:.5f}; "
          f"stats {st}", flush=True)
    assert st["failed_rows"] == 0
    users = degree_sample(ds.csr_row_ptr, 96, 6)
    items = degree_sample(ds.csc_col_ptr, 24, 7)
    for name, (rp, ci, cv, src, out, rows) in {
            "x": (ds.csr_row_ptr, ds.csr_col, ds.csr_val, bf16_round(T2), X3, users),
            "theta": (ds.csc_col_ptr, ds.csc_row, ds.csc_val, bf16_round(X3), T3, items)}.items():
        res, err = row_errors(rp, ci, cv, src, out, f, lam, rows)
        print(f"  {name}: residual {_q(res)}; factor error {_q(err)}", flush=True)
        # CG accepted at 1e-5 against the GPU's Gram, or Cholesky
        assert res.max() <= 2e-5, (name, res.max())
    print(f"zipf_4m_500k_1b: checked at {time.time() - t0:.1f}s")
