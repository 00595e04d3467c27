"""GPU checks of the mixed-precision solver (CUMF_SOLVER_CG) against the
FP64 oracle, through the C ABI.

The solver runs FP16 conjugate gradients on the Jacobi-scaled system
inside FP32 iterative refinement with an FP64 residual, and hands rows
whose relative residual stays above cg_tol to the Cholesky solve.  Bars:
  * every row: ||b - A x|| / ||b|| <= cg_tol (1 + 1e-3) against the Gram the
    GPU formed (evaluated here in FP64; 2 cg_tol for rows split across
    tasks, whose one-pass dump sums in another order than the partial
    Grams the solver used), and <= 2 cg_tol against the oracle's FP64
    system;
  * a tolerance FP32 cannot reach sends every row to the Cholesky solve,
    whose factors then match the Cholesky solver's to 1e-6;
  * full training: held-out RMSE within 1e-3 of the FP64 oracle.
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


def _warm_factors(ds, f, lam, iters, seed):
    T0 = cumf.init_factors(ds.n, f, seed)
    with _als(ds, f, lam) as h:
        h.set_factors(theta=T0)
        h.iterate(iters)
        return h.get_factors()


@pytest.mark.parametrize("f,tol", [(10, 1e-4), (20, 1e-4), (32, 1e-4), (100, 1e-4), (100, 1e-6), (120, 1e-4),
                                   (128, 1e-4)])
def test_cg_half_sweep_residuals(f, tol):
    ds = _data.netflix_like_small()
    lam = 0.058
    X0, T0 = _warm_factors(ds, f, lam, 2, 31)
    with _als(ds, f, lam, solver=G.CUMF_SOLVER_CG, cg_tol=tol, split=256) as h:
        h.set_factors(X=X0, theta=T0)
        h.update_x()
        X1, _ = h.get_factors()
        st = h.stats()
        rows = np.arange(0, ds.m, 13, dtype=np.int32)
        D = h.row_gram(0, rows).astype(np.float64)
    assert st["failed_rows"] == 0
    ldf = (f + 3) // 4 * 4
    counts = np.diff(ds.csr_row_ptr)
    worst_gpu = worst_split = worst_64 = 0.0
    for k, u in enumerate(rows):
        A, b = D[k, :f, :f], D[k, :f, ldf]
        x = X1[u].astype(np.float64)
        res = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
        if counts[u] <= 256:
            worst_gpu = max(worst_gpu, res)
        else:
            worst_split = max(worst_split, res)
        A64, b64 = O.row_system(ds.csr_row_ptr, ds.csr_col, ds.csr_val, T0, f, lam, u)
        worst_64 = max(worst_64, np.linalg.norm(b64 - A64 @ x) / np.linalg.norm(b64))
    # unsplit rows: the dumped Gram is the one the solver used, bit for bit
    assert worst_gpu <= tol * (1 + 1e-3), worst_gpu
    # split rows were solved on a sum of 256-rating partial Grams; the dump
    # forms the Gram in one pass (another FP32 summation order)
    assert worst_split <= 2 * tol, worst_split
    assert worst_64 <= 2 * tol, worst_64
    # split rows (heavy) went through the finish kernel's copy of the solver
    Xo, _ = O.half_sweep(ds.csr_row_ptr, ds.csr_col, ds.csr_val, T0, lam)
    rel = np.linalg.norm(X1 - Xo, axis=1) / np.maximum(np.linalg.norm(Xo, axis=1), 1e-30)
    assert np.median(rel) <= 100 * tol, np.median(rel)


@pytest.mark.parametrize("precision", [G.CUMF_PREC_FP32, G.CUMF_PREC_BF16_GRAM])
def test_cg_unreachable_tolerance_falls_back_to_cholesky(precision):
    ds = _data.netflix_like_small()
    f, lam = 100, 0.058
    X0, T0 = _warm_factors(ds, f, lam, 1, 32)
    with _als(ds, f, lam, precision=precision, split=256) as h:
        h.set_factors(X=X0, theta=T0)
        h.update_x()
        Xc, _ = h.get_factors()
    with _als(ds, f, lam, precision=precision, split=256, solver=G.CUMF_SOLVER_CG,
              cg_tol=1e-12, cg_rounds=1, cg_iters=1) as h:
        h.set_factors(X=X0, theta=T0)
        h.update_x()
        Xg, _ = h.get_factors()
        st = h.stats()
    assert st["cg_fallback_rows"] == ds.m
    assert st["failed_rows"] == 0
    rel = np.linalg.norm(Xg - Xc, axis=1) / np.maximum(np.linalg.norm(Xc, axis=1), 1e-30)
    assert rel.max() <= 1e-6, rel.max()


@pytest.mark.parametrize("precision", [G.CUMF_PREC_FP32, G.CUMF_PREC_BF16_GRAM])
def test_cg_training_rmse_vs_fp64_oracle(precision):
    """10,000 x 2,000 (450k training ratings), f=32, lambda=0.05, 10
    iterations from the same start: RMSE within 1e-3 of the FP64 oracle."""
    ds = _data.small_zipf()
    f, lam, iters = 32, 0.05, 10
    T0 = cumf.init_factors(ds.n, f, 0)
    _, _, hist = O.als_train(ds, f, lam, iters, T0)
    with _als(ds, f, lam, precision=precision, solver=G.CUMF_SOLVER_CG) as h:
        h.set_factors(theta=T0)
        h.iterate(iters)
        rm = h.rmse()
        st = h.stats()
    assert st["failed_rows"] == 0
    assert abs(rm - hist[-1]) <= 1e-3, (rm, hist[-1], st["cg_fallback_rows"])


def test_cg_training_f100_vs_fp64_oracle():
    """4,000 x 1,500 power-law matrix (404k ratings), f=100, lambda=0.058,
    5 iterations, BF16 Gram + CG: RMSE within 1e-3 of the FP64 oracle."""
    ds = _data.netflix_like_small()
    f, lam, iters = 100, 0.058, 5
    T0 = cumf.init_factors(ds.n, f, 5)
    _, _, hist = O.als_train(ds, f, lam, iters, T0)
    with _als(ds, f, lam, precision=G.CUMF_PREC_BF16_GRAM, solver=G.CUMF_SOLVER_CG) as h:
        h.set_factors(theta=T0)
        h.iterate(iters)
        rm = h.rmse()
    assert abs(rm - hist[-1]) <= 1e-3, (rm, hist[-1])


def test_cg_rejects_compat_mode():
    with pytest.raises(G.CumfError):
        G.ALS(100, 0.05, solver=G.CUMF_SOLVER_CG, compat_hermitian_bug=1)
