"""GPU parity of the HIP path against the FP64 oracle, through the C ABI.

Tolerances (FP32 Gram on MFMA + FP32 Cholesky vs an FP64 oracle):
  * Gram entries: |A_gpu - A_64| <= 2e-5 * max|diag A_64| per row (FP32
    accumulation of up to a few thousand terms).
  * Solve: relative residual of the GPU solution against the system the
    GPU formed, computed in FP64: ||b - A x|| / ||b|| <= 1e-5.
  * Solution vs oracle: ||x_gpu - x_64|| / ||x_64|| <= 1e-3 per row.
  * Held-out RMSE after full training: |RMSE_gpu - RMSE_64| <= 1e-3.
"""
import os

import numpy as np
import pytest

import cumf
from cumf import als as G
from oracle import oracle as O
from tests import _data

pytestmark = pytest.mark.gpu


def _als(ds, f, lam, **opt):
    h = G.ALS(f, lam, **opt)
    h.set_dataset(ds)
    return h


def test_mfma_layout():
    out = G.mfma_layout()
    A = 100.0 * np.arange(16)[:, None] + np.arange(4)[None, :]
    B = 1000.0 * np.arange(4)[:, None] + np.arange(16)[None, :]
    D = A @ B
    lane = np.arange(64)
    for r in range(4):
        np.testing.assert_array_equal(out[:, r], D[4 * (lane // 16) + r, lane % 16])


@pytest.mark.parametrize("f", [32, 100])
def test_gram_matches_oracle(f):
    ds = _data.netflix_like_small()
    lam = 0.058
    T0 = cumf.init_factors(ds.n, f, 1)
    with _als(ds, f, lam) as h:
        h.set_factors(theta=T0)
        deg = np.diff(ds.csr_row_ptr)
        rows = np.concatenate([np.argsort(-deg)[:4], np.argsort(deg)[:4], [7, 123, 999]]).astype(np.int32)
        D = h.row_gram(0, rows)
    ldf = (f + 3) // 4 * 4
    for k, u in enumerate(rows):
        A, b = O.row_system(ds.csr_row_ptr, ds.csr_col, ds.csr_val, T0, f, lam, u)
        scale = np.abs(np.diag(A)).max()
        assert np.abs(D[k, :f, :f] - A).max() <= 2e-5 * scale, u
        assert np.abs(D[k, :f, ldf] - b).max() <= 2e-5 * max(np.abs(b).max(), 1.0), u
        # padding columns stay zero
        assert np.all(D[k, :f, f:ldf] == 0)


def _check_sweep(ds_rp, ds_ci, ds_cv, theta, Xg, f, lam, rows):
    Xo, st = O.half_sweep(ds_rp, ds_ci, ds_cv, theta, lam)
    assert st.max() == 0
    rel = np.linalg.norm(Xg - Xo, axis=1) / np.maximum(np.linalg.norm(Xo, axis=1), 1e-30)
    assert rel.max() <= 1e-3, (rel.max(), int(rel.argmax()))
    for u in rows:
        A, b = O.row_system(ds_rp, ds_ci, ds_cv, theta, f, lam, u)
        res = np.linalg.norm(b - A @ Xg[u].astype(np.float64)) / np.linalg.norm(b)
        assert res <= 1e-4, (u, res)
    return rel


@pytest.mark.parametrize("f,split", [(32, 2048), (100, 2048), (100, 64), (10, 200), (128, 300)])
def test_half_sweeps_match_oracle(f, split):
    ds = _data.netflix_like_small()
    lam = 0.058
    T0 = cumf.init_factors(ds.n, f, 2)
    with _als(ds, f, lam, split=split) as h:
        h.set_factors(theta=T0)
        h.update_x()
        X1, _ = h.get_factors()
        h.update_theta()
        _, T1 = h.get_factors()
        assert h.stats()["failed_rows"] == 0
    _check_sweep(ds.csr_row_ptr, ds.csr_col, ds.csr_val, T0, X1, f, lam, [0, 1, 2])
    _check_sweep(ds.csc_col_ptr, ds.csc_row, ds.csc_val, X1, T1, f, lam, [0, 1, 2])


def test_solve_residual_against_gpu_system():
    """Relative residual of each GPU solve, against the Gram the GPU formed,
    in FP64: <= 1e-5."""
    ds = _data.netflix_like_small()
    f, lam = 100, 0.058
    T0 = cumf.init_factors(ds.n, f, 3)
    with _als(ds, f, lam) as h:
        h.set_factors(theta=T0)
        h.update_x()
        X1, _ = h.get_factors()
        rows = np.arange(0, ds.m, 97, dtype=np.int32)
        D = h.row_gram(0, rows).astype(np.float64)
    for k, u in enumerate(rows):
        A, b = D[k, :f, :f], D[k, :f, f]
        res = np.linalg.norm(b - A @ X1[u].astype(np.float64)) / np.linalg.norm(b)
        assert res <= 1e-5, (u, res)


def test_full_als_small_zipf_rmse():
    """m=10,000 n=2,000 (450k training ratings), f=32, lambda=0.05, 10
    iterations: held-out RMSE within 1e-3 of the FP64 oracle."""
    ds = _data.small_zipf()
    f, lam, iters = 32, 0.05, 10
    T0 = cumf.init_factors(ds.n, f, 0)
    _, _, hist = O.als_train(ds, f, lam, iters, T0)
    with _als(ds, f, lam) as h:
        h.set_factors(theta=T0)
        rm = []
        for _ in range(iters):
            h.iterate(1)
            rm.append(h.rmse())
    assert abs(rm[-1] - hist[-1]) <= 1e-3, (rm, hist)
    assert np.all(np.abs(np.array(rm) - hist) <= 1e-3)


def test_deterministic():
    ds = _data.netflix_like_small()
    f, lam = 100, 0.058
    T0 = cumf.init_factors(ds.n, f, 4)
    outs = []
    for _ in range(2):
        with _als(ds, f, lam, split=128) as h:
            h.set_factors(theta=T0)
            h.iterate(2)
            outs.append(h.get_factors())
    np.testing.assert_array_equal(outs[0][0], outs[1][0])
    np.testing.assert_array_equal(outs[0][1], outs[1][1])


def test_empty_rows_give_zero_factors():
    rp = np.array([0, 2, 2, 5, 5], np.int32)  # rows 1 and 3 empty
    ci = np.array([0, 2, 1, 2, 3], np.int32)
    cv = np.array([4, 5, 3, 1, 2], np.float32)
    f, lam = 8, 0.1
    T0 = cumf.init_factors(4, f, 5)
    with G.ALS(f, lam) as h:
        h.set_ratings(4, 4, rp, ci, cv)
        h.set_factors(theta=T0)
        h.update_x()
        X, _ = h.get_factors()
    Xo, _ = O.half_sweep(rp, ci, cv, T0, lam)
    assert np.all(X[[1, 3]] == 0)
    np.testing.assert_allclose(X, Xo, rtol=1e-4, atol=1e-5)


@pytest.mark.parametrize("f", [10, 100])
def test_do_als_with_csr_dropin(f):
    """The legacy entry point: CSR in mallocBest memory (consumed)."""
    ds = _data.netflix_like_small()
    lam = 0.058
    T0 = cumf.init_factors(ds.n, f, 6)
    X = G.do_als_with_csr(ds.csr_row_ptr, ds.csr_col, ds.csr_val, T0, ds.m, ds.n, f, lam)
    Xo, _ = O.half_sweep(ds.csr_row_ptr, ds.csr_col, ds.csr_val, T0, lam)
    rel = np.linalg.norm(X - Xo, axis=1) / np.linalg.norm(Xo, axis=1)
    assert rel.max() <= 1e-3


def test_do_als_with_csr_rejects_bad_f():
    ds = _data.tiny()
    T0 = cumf.init_factors(ds.n, 4, 0)
    with pytest.raises(G.CumfError):
        G.do_als_with_csr(ds.csr_row_ptr, ds.csr_col, ds.csr_val, T0, ds.m, ds.n, 0, 0.1)


def test_host_memory_fallback():
    """mallocBest's mapped-host fallback still gives the right answer."""
    ds = _data.tiny()
    f, lam = 16, 0.1
    T0 = cumf.init_factors(ds.n, f, 8)
    G.lib().cumf_debug_force_host_alloc(1)
    try:
        X = G.do_als_with_csr(ds.csr_row_ptr, ds.csr_col, ds.csr_val, T0, ds.m, ds.n, f, lam)
    finally:
        G.lib().cumf_debug_force_host_alloc(0)
    Xo, _ = O.half_sweep(ds.csr_row_ptr, ds.csr_col, ds.csr_val, T0, lam)
    np.testing.assert_allclose(X, Xo, rtol=1e-3, atol=1e-4)


def test_device_transpose_matches_host_csc():
    """Handles built with and without a host CSC (the latter transposed on
    the device) give bit-identical Theta sweeps."""
    ds = _data.netflix_like_small()
    f, lam = 32, 0.058
    T0 = cumf.init_factors(ds.n, f, 9)
    outs = []
    for with_csc in (True, False):
        with G.ALS(f, lam) as h:
            if with_csc:
                h.set_ratings(ds.m, ds.n, ds.csr_row_ptr, ds.csr_col, ds.csr_val,
                              ds.csc_col_ptr, ds.csc_row, ds.csc_val)
            else:
                h.set_ratings(ds.m, ds.n, ds.csr_row_ptr, ds.csr_col, ds.csr_val)
            h.set_factors(theta=T0)
            h.iterate(1)
            outs.append(h.get_factors())
    np.testing.assert_array_equal(outs[0][0], outs[1][0])
    np.testing.assert_array_equal(outs[0][1], outs[1][1])


def test_concurrent_dropin_callers():
    """Several host threads call doALSWithCSR at once (Spark executors share
    a GPU): every result matches the oracle."""
    import threading
    ds = _data.netflix_like_small()
    f, lam = 16, 0.058
    results, errors = {}, []

    def work(i):
        try:
            T0 = cumf.init_factors(ds.n, f, 100 + i)
            results[i] = (T0, G.do_als_with_csr(ds.csr_row_ptr, ds.csr_col, ds.csr_val, T0, ds.m, ds.n, f, lam))
        except Exception as e:  # noqa: BLE001
            errors.append(e)

    threads = [threading.Thread(target=work, args=(i,)) for i in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    for i, (T0, X) in results.items():
        Xo, _ = O.half_sweep(ds.csr_row_ptr, ds.csr_col, ds.csr_val, T0, lam)
        rel = np.linalg.norm(X - Xo, axis=1) / np.linalg.norm(Xo, axis=1)
        assert rel.max() <= 1e-3, i


def test_runals_driver_small_preset():
    import json
    import subprocess
    from cumf import _paths
    exe = os.path.join(_paths.PKG_DIR, "runals")
    r = subprocess.run([exe, "gpu", "32", "3", "0.05", "zipf_10k_2k_500k"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    summary = json.loads(r.stdout.strip().splitlines()[-1])
    assert summary["iterations"] == 3 and summary["failed_rows"] == 0
    assert 0.2 < summary["heldout_rmse"] < 1.0


def test_runals_mixed_path_with_autotune(tmp_path):
    import json
    import subprocess
    from cumf import _paths
    exe = os.path.join(_paths.PKG_DIR, "runals")
    table = str(tmp_path / "tiles.txt")
    r = subprocess.run([exe, "gpu", "32", "3", "0.05", "zipf_10k_2k_500k", "--precision", "bf16", "--solver", "cg",
                        "--autotune", table, "--graph", "1"], capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stderr
    summary = json.loads(r.stdout.strip().splitlines()[-1])
    assert summary["precision"] == "bf16" and summary["solver"] == "cg" and summary["failed_rows"] == 0
    assert 0.2 < summary["heldout_rmse"] < 1.0
    r2 = subprocess.run([exe, "gpu", "32", "3", "0.05", "zipf_10k_2k_500k", "--precision", "bf16", "--solver", "cg",
                         "--tiles", table], capture_output=True, text=True, timeout=180)
    assert r2.returncode == 0, r2.stderr
    assert abs(json.loads(r2.stdout.strip().splitlines()[-1])["heldout_rmse"] - summary["heldout_rmse"]) < 1e-6


def test_runals_reads_reference_binary_csr(tmp_path):
    ds = _data.tiny()
    import subprocess
    from cumf import _paths
    pre = str(tmp_path / "r")
    G.save_csr_bin(pre, ds.csr_row_ptr, ds.csr_col, ds.csr_val)
    exe = os.path.join(_paths.PKG_DIR, "runals")
    r = subprocess.run([exe, "gpu", "8", "2", "0.1", f"csr:{ds.m}:{ds.n}:{ds.nnz}:{pre}"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert '"failed_rows": 0' in r.stdout


def test_graph_replay_matches_eager():
    """use_graph: one captured graph per half-sweep, replayed several times
    with eager work between replays, equals eager iterations bit for bit."""
    ds = _data.netflix_like_small()
    f, lam = 32, 0.058
    T0 = cumf.init_factors(ds.n, f, 11)
    with _als(ds, f, lam, split=256) as eager, _als(ds, f, lam, split=256, use_graph=1, timing=1) as graph:
        eager.set_factors(theta=T0)
        graph.set_factors(theta=T0)
        for rep in range(3):
            eager.iterate(1)
            graph.iterate(1)
            graph.rmse()  # eager kernels on the same stream between replays
            Xe, Te = eager.get_factors()
            Xg, Tg = graph.get_factors()
            np.testing.assert_array_equal(Xe, Xg, err_msg=f"replay {rep}")
            np.testing.assert_array_equal(Te, Tg, err_msg=f"replay {rep}")
        st = graph.stats()
        assert st["iters"] == 6 and st["iter_ms"] > 0  # two half-sweep replays per iteration


@pytest.mark.parametrize("f", [10, 100])
def test_compat_system_bit_exact(f):
    """compat_hermitian_bug: the system the GPU solver sees equals, bit for
    bit, the oracle's restatement of the reference's FP32 arithmetic
    (sequential FMAs in rating order, the (10I+4, 10J) slip, FP32 ridge)."""
    ds = _data.netflix_like_small()
    lam = 0.058
    T0 = cumf.init_factors(ds.n, f, 12)
    X0 = cumf.init_factors(ds.m, f, 13)
    with _als(ds, f, lam, compat_hermitian_bug=1) as h:
        h.set_factors(X=X0, theta=T0)
        deg = np.diff(ds.csr_row_ptr)
        rows = np.array([0, 5, int(np.argmax(deg)), int(np.argmin(deg))], np.int32)
        M, b = h.compat_system(0, rows)
        for k, u in enumerate(rows):
            Mo, bo = O.compat_system(ds.csr_row_ptr, ds.csr_col, ds.csr_val, T0, f, lam, u)
            np.testing.assert_array_equal(M[k], Mo)
            np.testing.assert_array_equal(b[k], bo)
        cols = np.array([0, 7], np.int32)
        M, b = h.compat_system(1, cols)
        for k, v in enumerate(cols):
            Mo, bo = O.compat_system(ds.csc_col_ptr, ds.csc_row, ds.csc_val, X0, f, lam, v)
            np.testing.assert_array_equal(M[k], Mo)
            np.testing.assert_array_equal(b[k], bo)


@pytest.mark.parametrize("f", [10, 100])
def test_compat_half_sweep_matches_oracle(f):
    """Reference-reproducing factors: LU (no pivoting) of the slipped system,
    against the oracle's FP64 LU of the same system, and within FP32 Gram
    rounding of the FP64 compat half-sweep."""
    ds = _data.netflix_like_small()
    lam = 0.058
    T0 = cumf.init_factors(ds.n, f, 14)
    with _als(ds, f, lam, compat_hermitian_bug=1) as h:
        h.set_factors(theta=T0)
        h.update_x()
        X, _ = h.get_factors()
        assert h.stats()["failed_rows"] == 0
    for u in (0, 1, 2, 99):
        Mo, bo = O.compat_system(ds.csr_row_ptr, ds.csr_col, ds.csr_val, T0, f, lam, u)
        np.testing.assert_allclose(X[u], O.lu_solve(Mo, bo), rtol=1e-5, atol=1e-6)
    Xo, _ = O.half_sweep(ds.csr_row_ptr, ds.csr_col, ds.csr_val, T0, lam, compat=True)
    rel = np.linalg.norm(X - Xo, axis=1) / np.linalg.norm(Xo, axis=1)
    assert rel.max() <= 1e-3
    # and it is NOT the clean update
    Xc, _ = O.half_sweep(ds.csr_row_ptr, ds.csr_col, ds.csr_val, T0, lam)
    assert (np.linalg.norm(X - Xc, axis=1) / np.linalg.norm(Xc, axis=1)).max() > 1e-3


def test_compat_env_switch_on_dropin(monkeypatch):
    ds = _data.netflix_like_small()
    f, lam = 10, 0.058
    T0 = cumf.init_factors(ds.n, f, 15)
    monkeypatch.setenv("CUMF_COMPAT_HERMITIAN_BUG", "1")
    X = G.do_als_with_csr(ds.csr_row_ptr, ds.csr_col, ds.csr_val, T0, ds.m, ds.n, f, lam)
    monkeypatch.delenv("CUMF_COMPAT_HERMITIAN_BUG")
    Xo, _ = O.half_sweep(ds.csr_row_ptr, ds.csr_col, ds.csr_val, T0, lam, compat=True)
    rel = np.linalg.norm(X - Xo, axis=1) / np.linalg.norm(Xo, axis=1)
    assert rel.max() <= 1e-3


def _bf16(x):
    """float32 values rounded to bfloat16 (round to nearest even), as float32."""
    b = np.ascontiguousarray(x, np.float32).view(np.uint32)
    b = (b + (((b >> 16) & 1) + 0x7FFF)) & 0xFFFF0000
    return b.view(np.float32)


def test_mfma_bf16_layout():
    out = G.mfma_bf16_layout()
    k = np.arange(32)
    A = np.arange(16)[:, None] + 2 * (k[None, :] % 4)
    B = (k[:, None] % 5) + 3 * np.arange(16)[None, :]
    D = (A @ B).astype(np.float64)
    lane = np.arange(64)
    for r in range(4):
        np.testing.assert_array_equal(out[:, r], D[4 * (lane // 16) + r, lane % 16])


@pytest.mark.parametrize("f", [32, 100])
def test_bf16_gram_matches_oracle_on_rounded_factors(f):
    """CUMF_PREC_BF16_GRAM: the Gram equals the FP64 Gram of the BF16-rounded
    factors to FP32 accumulation error; b (hi + lo columns) likewise."""
    ds = _data.netflix_like_small()
    lam = 0.058
    T0 = cumf.init_factors(ds.n, f, 21)
    with _als(ds, f, lam, precision=G.CUMF_PREC_BF16_GRAM) as h:
        h.set_factors(theta=T0)
        deg = np.diff(ds.csr_row_ptr)
        rows = np.concatenate([np.argsort(-deg)[:3], np.argsort(deg)[:3], [7]]).astype(np.int32)
        D = h.row_gram(0, rows)
    Tr = _bf16(T0)
    for k, u in enumerate(rows):
        A, b = O.row_system(ds.csr_row_ptr, ds.csr_col, ds.csr_val, Tr, f, lam, u)
        scale = np.abs(np.diag(A)).max()
        assert np.abs(D[k, :f, :f] - A).max() <= 2e-5 * scale, u
        bg = D[k, :f, f].astype(np.float64) + D[k, :f, f + 1]
        assert np.abs(bg - b).max() <= 2e-5 * max(np.abs(b).max(), 1.0), u


def test_bf16_half_sweep_matches_oracle_on_rounded_factors():
    ds = _data.netflix_like_small()
    f, lam = 100, 0.058
    T0 = cumf.init_factors(ds.n, f, 22)
    with _als(ds, f, lam, precision=G.CUMF_PREC_BF16_GRAM, split=256) as h:
        h.set_factors(theta=T0)
        h.update_x()
        X1, _ = h.get_factors()
        assert h.stats()["failed_rows"] == 0
    Xo, _ = O.half_sweep(ds.csr_row_ptr, ds.csr_col, ds.csr_val, _bf16(T0), lam)
    rel = np.linalg.norm(X1 - Xo, axis=1) / np.linalg.norm(Xo, axis=1)
    assert rel.max() <= 1e-3


def test_bf16_gram_training_rmse_vs_fp64_oracle():
    """Full training (10,000 x 2,000, f=32, lambda=0.05, 10 iterations) with
    the BF16 Gram: held-out RMSE within 1e-3 of the (unrounded) FP64 oracle."""
    ds = _data.small_zipf()
    f, lam, iters = 32, 0.05, 10
    T0 = cumf.init_factors(ds.n, f, 0)
    _, _, hist = O.als_train(ds, f, lam, iters, T0)
    with _als(ds, f, lam, precision=G.CUMF_PREC_BF16_GRAM) as h:
        h.set_factors(theta=T0)
        h.iterate(iters)
        rm = h.rmse()
    assert abs(rm - hist[-1]) <= 1e-3, (rm, hist[-1])
