"""GPU parity of the HIP path against the FP64 oracle, through the C ABI.

Tolerances (FP32 Gram on MFMA + FP32 Cholesky vs an FP64 oracle):
  * Gram entries: |A_gpu - A_64| <= 2e-5 * max|diag A_64| per row (FP32
    accumulation of up to a few thousand terms).
  * Solve: relative residual of the GPU solution against the system the
    GPU formed, computed in FP64: ||b - A x|| / ||b|| <= 1e-5.
  * Solution vs oracle: ||x_gpu - x_64|| / ||x_64|| <= 1e-3 per row.
  * Held-out RMSE after full training: |RMSE_gpu - RMSE_64| <= 1e-3.
"""
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


@pytest.mark.parametrize("f,split", [(32, 2048), (100, 2048), (100, 64), (10, 200)])
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
