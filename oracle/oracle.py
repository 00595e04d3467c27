"""ctypes binding of the FP64 CPU oracle, oracle/libals_oracle.so.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
the CPU-baseline leg of bench.py, always as the oracle or the timed CPU
baseline, never as part of the product path.  See als_oracle.c for the
reference lines each function restates.
"""
import ctypes as C
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libals_oracle.so")
_lib = None
P = C.c_void_p
I64 = C.c_int64


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(LIB_PATH)
        L.oracle_max_threads.restype = C.c_int
        L.oracle_row_system.argtypes = [P, P, P, P, I64, C.c_int, C.c_float, I64, C.c_int, P, P]
        L.oracle_chol_solve.argtypes = [C.c_int, P, P]
        L.oracle_compat_system.argtypes = [P, P, P, P, C.c_int, C.c_float, I64, P, P]
        L.oracle_lu_solve.argtypes = [C.c_int, P, P]
        L.oracle_half_sweep_d.argtypes = [I64, I64, C.c_int, P, P, P, P, C.c_float, C.c_int, P, P, C.c_int]
        L.oracle_half_sweep_f.argtypes = [I64, I64, C.c_int, P, P, P, P, C.c_float, C.c_int, P, P, C.c_int]
        L.oracle_rmse_d.argtypes = [I64, P, P, P, C.c_int, P, P, C.c_int]
        L.oracle_rmse_d.restype = C.c_double
        L.oracle_rmse_f.argtypes = [I64, P, P, P, C.c_int, P, P, C.c_int]
        L.oracle_rmse_f.restype = C.c_double
        L.oracle_objective_d.argtypes = [I64, I64, C.c_int, P, P, P, P, P, C.c_float, C.c_int]
        L.oracle_objective_d.restype = C.c_double
        L.oracle_als_train.argtypes = [I64, I64, C.c_int, P, P, P, P, P, P, P, C.c_float, C.c_int,
                                       C.c_int, I64, P, P, P, P, P, P, C.c_int]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(P)


def _c(a, t):
    return np.ascontiguousarray(a, t)


def max_threads():
    return lib().oracle_max_threads()


def row_system(row_ptr, col, val, theta, f, lam, u, compat=False):
    """(A, b) of row u in FP64: A = Y^T Y + lam n_u I, b = Y^T r."""
    row_ptr, col, val = _c(row_ptr, np.int32), _c(col, np.int32), _c(val, np.float32)
    theta = _c(theta, np.float32)
    n = theta.shape[0]
    A = np.empty((f, f), np.float64)
    b = np.empty(f, np.float64)
    rc = lib().oracle_row_system(_p(row_ptr), _p(col), _p(val), _p(theta), n, f, lam, int(u),
                                 int(compat), _p(A), _p(b))
    if rc:
        raise ValueError(f"oracle_row_system: {rc}")
    return A, b


def compat_system(row_ptr, col, val, theta, f, lam, u):
    """(M, b) the reference's solver sees for row u: FP32 fused multiply-adds
    in rating order, the Gram write slip, the FP32 ridge (FP64 out)."""
    row_ptr, col, val = _c(row_ptr, np.int32), _c(col, np.int32), _c(val, np.float32)
    theta = _c(theta, np.float32)
    M = np.empty((f, f), np.float64)
    b = np.empty(f, np.float64)
    rc = lib().oracle_compat_system(_p(row_ptr), _p(col), _p(val), _p(theta), f, lam, int(u), _p(M), _p(b))
    if rc:
        raise ValueError(f"oracle_compat_system: {rc}")
    return M, b


def chol_solve(A, b):
    A = np.array(A, np.float64, order="C", copy=True)
    x = np.array(b, np.float64, copy=True)
    rc = lib().oracle_chol_solve(A.shape[0], _p(A), _p(x))
    if rc:
        raise np.linalg.LinAlgError(f"oracle_chol_solve: {rc}")
    return x


def lu_solve(A, b):
    A = np.array(A, np.float64, order="C", copy=True)
    x = np.array(b, np.float64, copy=True)
    rc = lib().oracle_lu_solve(A.shape[0], _p(A), _p(x))
    if rc:
        raise np.linalg.LinAlgError(f"oracle_lu_solve: {rc}")
    return x


def half_sweep(row_ptr, col, val, theta, lam, compat=False, threads=0):
    """x_u for every row of the CSR from factors theta (n x f).

    theta float32 -> returns float32 X (FP64 internally); theta float64 ->
    returns float64 X.  Also returns the per-row solver status.
    """
    row_ptr, col, val = _c(row_ptr, np.int32), _c(col, np.int32), _c(val, np.float32)
    m = row_ptr.shape[0] - 1
    n, f = theta.shape
    status = np.zeros(m, np.int32)
    if theta.dtype == np.float64:
        theta = _c(theta, np.float64)
        X = np.empty((m, f), np.float64)
        rc = lib().oracle_half_sweep_d(m, n, f, _p(row_ptr), _p(col), _p(val), _p(theta), lam,
                                       int(compat), _p(X), _p(status), threads)
    else:
        theta = _c(theta, np.float32)
        X = np.empty((m, f), np.float32)
        rc = lib().oracle_half_sweep_f(m, n, f, _p(row_ptr), _p(col), _p(val), _p(theta), lam,
                                       int(compat), _p(X), _p(status), threads)
    if rc:
        raise ValueError(f"oracle half sweep: {rc}")
    return X, status


def rmse(u, v, r, X, T, threads=0):
    u, v, r = _c(u, np.int32), _c(v, np.int32), _c(r, np.float32)
    f = X.shape[1]
    if X.dtype == np.float64:
        return lib().oracle_rmse_d(u.shape[0], _p(u), _p(v), _p(r), f, _p(_c(X, np.float64)),
                                   _p(_c(T, np.float64)), threads)
    return lib().oracle_rmse_f(u.shape[0], _p(u), _p(v), _p(r), f, _p(_c(X, np.float32)),
                               _p(_c(T, np.float32)), threads)


def objective(row_ptr, col, val, X, T, lam, threads=0):
    row_ptr, col, val = _c(row_ptr, np.int32), _c(col, np.int32), _c(val, np.float32)
    X, T = _c(X, np.float64), _c(T, np.float64)
    return lib().oracle_objective_d(X.shape[0], T.shape[0], X.shape[1], _p(row_ptr), _p(col), _p(val),
                                    _p(X), _p(T), lam, threads)


def als_train(ds, f, lam, iters, theta0, compat=False, threads=0):
    """FP64 ALS (X sweep then Theta sweep per iteration) on a cumf Dataset.

    Returns (X, T, rmse_per_iteration) with FP64 factors."""
    theta0 = _c(theta0, np.float32)
    X = np.empty((ds.m, f), np.float64)
    T = np.empty((ds.n, f), np.float64)
    hist = np.zeros(iters, np.float64)
    rc = lib().oracle_als_train(ds.m, ds.n, f, _p(ds.csr_row_ptr), _p(ds.csr_col), _p(ds.csr_val),
                                _p(ds.csc_col_ptr), _p(ds.csc_row), _p(ds.csc_val), _p(theta0), lam,
                                iters, int(compat), ds.test_u.shape[0], _p(ds.test_u), _p(ds.test_v),
                                _p(ds.test_r), _p(X), _p(T), _p(hist), threads)
    if rc:
        raise ValueError(f"oracle_als_train: {rc}")
    return X, T, hist
