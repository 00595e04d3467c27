"""The host ALS half-sweep of libcumf_cpu.so (the standalone driver's cpu
mode) against the golden fixtures and the FP64 oracle; CPU only."""
import ctypes as C
import os

import numpy as np

import cumf
from cumf import _paths
from oracle import oracle as O
from tests import _data

LIB = C.CDLL(_paths.CPU_LIB)
P = C.c_void_p
LIB.cumf_cpu_half_sweep.argtypes = [C.c_int64, C.c_int64, C.c_int, P, P, P, P, C.c_float, P, C.c_int]
LIB.cumf_cpu_half_sweep.restype = C.c_int
LIB.cumf_cpu_transpose.argtypes = [C.c_int64, C.c_int64, C.c_int64, P, P, P, P, P, P]
LIB.cumf_cpu_transpose.restype = C.c_int
LIB.cumf_cpu_rmse.argtypes = [C.c_int64, P, P, P, C.c_int, P, P, C.c_int]
LIB.cumf_cpu_rmse.restype = C.c_double
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "als_fixtures.npz")


def _p(a):
    return a.ctypes.data_as(P)


def half_sweep(rp, ci, cv, theta, lam, threads=4):
    rp, ci, cv = (np.ascontiguousarray(a, t) for a, t in ((rp, np.int32), (ci, np.int32), (cv, np.float32)))
    theta = np.ascontiguousarray(theta, np.float32)
    m, (n, f) = rp.shape[0] - 1, theta.shape
    X = np.full((m, f), np.nan, np.float32)
    rc = LIB.cumf_cpu_half_sweep(m, n, f, _p(rp), _p(ci), _p(cv), _p(theta), lam, _p(X), threads)
    return X, rc


def test_half_sweep_matches_golden_vectors():
    """X1 of the fixtures: dense numpy/scipy half-sweep (FP64) from theta;
    the host sweep stores FP32, so 1e-6 relative."""
    g = np.load(GOLDEN)
    for f in (1, 4, 10, 32, 100):
        X, rc = half_sweep(g["rp"], g["ci"], g["cv"], g[f"theta_f{f}"], float(g["lam"]))
        assert rc == 0
        ref = g[f"X1_f{f}"]
        rel = np.linalg.norm(X - ref, axis=1) / np.maximum(np.linalg.norm(ref, axis=1), 1e-30)
        assert rel.max() <= 1e-6, (f, rel.max())


def test_half_sweep_matches_oracle_and_empty_rows():
    ds = _data.netflix_like_small()
    f, lam = 32, 0.058
    T = cumf.init_factors(ds.n, f, 5)
    X, rc = half_sweep(ds.csr_row_ptr, ds.csr_col, ds.csr_val, T, lam)
    Xo, st = O.half_sweep(ds.csr_row_ptr, ds.csr_col, ds.csr_val, T, lam)
    assert rc == 0 and st.max() == 0
    rel = np.linalg.norm(X - Xo, axis=1) / np.maximum(np.linalg.norm(Xo, axis=1), 1e-30)
    assert rel.max() <= 1e-5, rel.max()
    rp = np.array([0, 2, 2, 3], np.int32)  # row 1 has no ratings
    X, rc = half_sweep(rp, np.array([0, 1, 1], np.int32), np.array([3, 4, 5], np.float32), T[:2, :4].copy(), lam)
    assert rc == 0 and np.all(X[1] == 0) and np.all(np.isfinite(X))


def test_transpose_and_rmse():
    ds = _data.tiny()
    cp = np.empty(ds.n + 1, np.int32)
    ri = np.empty(ds.nnz, np.int32)
    cv = np.empty(ds.nnz, np.float32)
    assert LIB.cumf_cpu_transpose(ds.m, ds.n, ds.nnz, _p(ds.csr_row_ptr), _p(ds.csr_col), _p(ds.csr_val),
                                  _p(cp), _p(ri), _p(cv)) == 0
    np.testing.assert_array_equal(cp, ds.csc_col_ptr)
    np.testing.assert_array_equal(ri, ds.csc_row)
    np.testing.assert_array_equal(cv, ds.csc_val)
    f = 8
    X = cumf.init_factors(ds.m, f, 1)
    T = cumf.init_factors(ds.n, f, 2)
    got = LIB.cumf_cpu_rmse(ds.test_u.shape[0], _p(ds.test_u), _p(ds.test_v), _p(ds.test_r), f, _p(X), _p(T), 2)
    pred = np.einsum("ij,ij->i", X[ds.test_u].astype(np.float64), T[ds.test_v].astype(np.float64))
    assert abs(got - np.sqrt(np.mean((ds.test_r - pred) ** 2))) <= 1e-12


def test_rejects_bad_input():
    T = np.ones((3, 2), np.float32)
    _, rc = half_sweep(np.array([0, 1], np.int32), np.array([7], np.int32), np.ones(1, np.float32), T, 0.1)
    assert rc == -1
    _, rc = half_sweep(np.array([0, 2, 1], np.int32), np.array([0, 1], np.int32), np.ones(2, np.float32), T, 0.1)
    assert rc == -1
