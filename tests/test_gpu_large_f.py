"""GPU checks of the half-sweep for 140 < f <= 256 (one work-group per row,
FP32 Gram + FP32 Cholesky in LDS; the dual form for lighter rows: up to 95
ratings on one wavefront, above on the work-group while the dual system
is the smaller), through the C ABI, against the FP64 oracle:
factors within 1e-3 relative, per-row relative residual against the
oracle's FP64 system <= 1e-4 (<= 1e-5 for the dual-form rows), the dual
form against the primal kernel on the same rows, empty rows, the drop-in
entry point (also with f % 4 != 0, the re-strided path), a short training
run's held-out RMSE, and the options that path does not offer are
refused."""
import functools

import numpy as np
import pytest

import cumf
from cumf import als as G
from oracle import oracle as O
from tests import _data

pytestmark = pytest.mark.gpu


@functools.lru_cache(maxsize=None)
def _mid():
    """m=800, n=400, 60,000 ratings: rows of 1..400 ratings."""
    return cumf.generate(_data.params(800, 400, 60600, 600, row_cap=400, col_cap=800, seed=3))


def _sweep_check(rp, ci, cv, theta, Xg, f, lam, rows):
    Xo, st = O.half_sweep(rp, ci, cv, theta, lam)
    assert st.max() == 0
    rel = np.linalg.norm(Xg - Xo, axis=1) / np.maximum(np.linalg.norm(Xo, axis=1), 1e-30)
    assert rel.max() <= 1e-3, (rel.max(), int(rel.argmax()))
    for u in rows:
        A, b = O.row_system(rp, ci, cv, theta, f, lam, u)
        res = np.linalg.norm(b - A @ Xg[u].astype(np.float64)) / np.linalg.norm(b)
        assert res <= 1e-4, (u, res)


@pytest.mark.parametrize("f", [141, 160, 256])
def test_large_f_half_sweeps_match_oracle(f):
    ds = _mid()
    lam = 0.058
    T0 = cumf.init_factors(ds.n, f, 11)
    with G.ALS(f, lam) as h:
        h.set_dataset(ds)
        h.set_factors(theta=T0)
        h.update_x()
        X1, _ = h.get_factors()
        h.update_theta()
        _, T1 = h.get_factors()
        assert h.stats()["failed_rows"] == 0
    _sweep_check(ds.csr_row_ptr, ds.csr_col, ds.csr_val, T0, X1, f, lam, [0, 1, 2, 399])
    _sweep_check(ds.csc_col_ptr, ds.csc_row, ds.csc_val, X1, T1, f, lam, [0, 1, 2])


def _x_sweep(ds, f, lam, T0, dual):
    prev = G.lib().cumf_debug_large_dual(int(dual))
    try:
        with G.ALS(f, lam) as h:
            h.set_dataset(ds)
            h.set_factors(theta=T0)
            h.update_x()
            X, _ = h.get_factors()
            assert h.stats()["failed_rows"] == 0
    finally:
        G.lib().cumf_debug_large_dual(prev)
    return X


def _sweep_nb(f):
    ldf, ldb = (f + 3) & ~3, (f + 7) & ~7
    return (max(ldf + 1, ldb, f + 2) + 15) // 16


@pytest.mark.parametrize("f", [144, 256])
def test_large_f_dual_form_matches_primal_and_oracle(f):
    """The dual forms against the primal kernel (switched on for every row)
    and the oracle's FP64 system: rows of at most 95 ratings on one
    wavefront (K = Y Y^T on FP32 MFMA, at most six blocks), rows whose dual
    system has at most NB - 2 blocks on the work-group; heavier rows are
    the primal kernel's either way, bit for bit."""
    ds = _mid()
    lam = 0.058
    T0 = cumf.init_factors(ds.n, f, 21)
    Xp = _x_sweep(ds, f, lam, T0, dual=False)
    Xd = _x_sweep(ds, f, lam, T0, dual=True)
    deg = np.diff(ds.csr_row_ptr)
    thr = 16 * min(_sweep_nb(f) - 2, 16) - 1  # longest row of the work-group dual form
    light = np.nonzero(deg <= 95)[0]
    medium = np.nonzero((deg > 95) & (deg <= thr))[0]
    heavy = np.nonzero(deg > thr)[0]
    assert len(light) >= 100 and len(medium) >= 10 and len(heavy) >= 5, (len(light), len(medium), len(heavy))
    np.testing.assert_array_equal(Xd[heavy], Xp[heavy])
    for rows in (light, medium):
        rel = np.linalg.norm(Xd[rows] - Xp[rows], axis=1) / np.linalg.norm(Xp[rows], axis=1)
        assert rel.max() <= 1e-4, (rel.max(), int(rows[rel.argmax()]))
    picks = set(light[:: max(1, len(light) // 32)]) | set(medium[:: max(1, len(medium) // 24)])
    for rows in (light, medium):
        picks |= {int(rows[deg[rows].argmin()]), int(rows[deg[rows].argmax()])}
    for u in sorted(picks):
        A, b = O.row_system(ds.csr_row_ptr, ds.csr_col, ds.csr_val, T0, f, lam, u)
        res = np.linalg.norm(b - A @ Xd[u].astype(np.float64)) / np.linalg.norm(b)
        assert res <= 1e-5, (u, int(deg[u]), res)


def test_large_f_empty_and_light_rows():
    """Empty rows give x = 0; rows at the block-count edges of both dual
    forms (1 .. 95 ratings on a wavefront, 96 .. 143 on the work-group at
    f = 160) and a heavier one, against the oracle."""
    f, lam, n = 160, 0.1, 200
    rng = np.random.default_rng(4)
    counts = [0, 1, 15, 16, 31, 0, 47, 63, 64, 79, 95, 96, 111, 112, 127, 143, 150]
    rp = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    ci = np.concatenate([np.sort(rng.choice(n, c, replace=False)) for c in counts]).astype(np.int32)
    cv = rng.integers(1, 6, size=ci.shape[0]).astype(np.float32)
    T0 = cumf.init_factors(n, f, 7)
    m = len(counts)
    with G.ALS(f, lam) as h:
        h.set_ratings(m, n, rp, ci, cv)
        h.set_factors(theta=T0)
        h.update_x()
        X, _ = h.get_factors()
        assert h.stats()["failed_rows"] == 0
    assert np.all(X[[0, 5]] == 0)
    for u in range(m):
        if counts[u] == 0:
            continue
        A, b = O.row_system(rp, ci, cv, T0, f, lam, u)
        res = np.linalg.norm(b - A @ X[u].astype(np.float64)) / np.linalg.norm(b)
        assert res <= 1e-5, (u, counts[u], res)


@pytest.mark.parametrize("f", [150, 200])
def test_large_f_dropin(f):
    """doALSWithCSR at f > 140; f = 150 is not a multiple of 4 (re-strided
    factor buffers)."""
    ds = _mid()
    lam = 0.05
    T0 = cumf.init_factors(ds.n, f, 12)
    X = G.do_als_with_csr(ds.csr_row_ptr, ds.csr_col, ds.csr_val, T0, ds.m, ds.n, f, lam)
    Xo, _ = O.half_sweep(ds.csr_row_ptr, ds.csr_col, ds.csr_val, T0, lam)
    rel = np.linalg.norm(X - Xo, axis=1) / np.linalg.norm(Xo, axis=1)
    assert rel.max() <= 1e-3, rel.max()


def test_large_f_training_rmse_vs_fp64_oracle():
    ds = _mid()
    f, lam, iters = 160, 0.1, 3
    T0 = cumf.init_factors(ds.n, f, 13)
    _, _, hist = O.als_train(ds, f, lam, iters, T0)
    with G.ALS(f, lam) as h:
        h.set_dataset(ds)
        h.set_factors(theta=T0)
        h.iterate(iters)
        rm = h.rmse()
    assert abs(rm - hist[-1]) <= 1e-3, (rm, hist[-1])


def test_large_f_refuses_other_paths():
    ds = _data.tiny()
    for opt in ({"solver": G.CUMF_SOLVER_CG}, {"precision": G.CUMF_PREC_BF16_GRAM},
                {"compat_hermitian_bug": 1}):
        with pytest.raises(G.CumfError):
            G.ALS(200, 0.05, **opt)
    with G.ALS(200, 0.05) as h:
        h.set_dataset(ds)
        h.set_factors(theta=cumf.init_factors(ds.n, 200, 0))
        with pytest.raises(G.CumfError):
            h.autotune(reps=1)
    with pytest.raises(G.CumfError):
        G.ALS(257, 0.05)
