"""Shared checks of GPU factors against the FP64 oracle (test helpers)."""
import concurrent.futures as cf
import os
import zlib

import numpy as np

from oracle import oracle as O

# Arrays of a generated dataset, in the order their CRC-32s are stored.
DATASET_ARRAYS = ("csr_row_ptr", "csr_col", "csr_val", "csc_col_ptr", "csc_row", "csc_val",
                  "test_u", "test_v", "test_r")


def dataset_crc(ds):
    """CRC-32 of every array of a generated dataset (int64 vector)."""
    return np.array([zlib.crc32(np.ascontiguousarray(getattr(ds, k)).view(np.uint8)) for k in DATASET_ARRAYS],
                    np.int64)


def bf16_round(a):
    """FP32 -> BF16 (round to nearest even) -> FP32: the factor copy the BF16
    Gram reads."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32).view(np.float32)


def degree_sample(row_ptr, k, seed):
    """k rows spread over the degree range: the heaviest eighth, the lightest
    eighth and a seeded uniform draw from the rest (rows without ratings
    excluded)."""
    deg = np.diff(row_ptr)
    live = np.flatnonzero(deg > 0)
    order = live[np.argsort(-deg[live], kind="stable")]
    pick = list(order[:k // 8]) + list(order[-(k // 8):])
    rest = np.setdiff1d(live, pick)
    pick += list(np.random.default_rng(seed).choice(rest, k - len(pick), replace=False))
    return np.sort(np.array(pick, np.int64))


def row_errors(row_ptr, col, val, src, X, f, lam, rows, threads=None):
    """Per sampled row u: relative residual ||b - A x_u|| / ||b|| of the GPU's
    x_u against the FP64 system (A, b) the oracle forms from the factors
    `src`, and the relative factor error ||x_u - x*|| / ||x*|| against the
    FP64 solution x* of that system.  Rows run on a thread pool (the
    oracle's C calls release the interpreter lock)."""
    def one(u):
        A, b = O.row_system(row_ptr, col, val, src, f, lam, int(u))
        x = X[u].astype(np.float64)
        xs = np.linalg.solve(A, b)
        return (np.linalg.norm(b - A @ x) / np.linalg.norm(b), np.linalg.norm(x - xs) / np.linalg.norm(xs))

    workers = threads or min(16, os.cpu_count() or 1)
    with cf.ThreadPoolExecutor(workers) as ex:
        out = list(ex.map(one, rows))
    return np.array([o[0] for o in out]), np.array([o[1] for o in out])


def rel_rows(A, B):
    """Row-wise ||a - b|| / ||b||."""
    A = np.asarray(A, np.float64)
    B = np.asarray(B, np.float64)
    return np.linalg.norm(A - B, axis=1) / np.maximum(np.linalg.norm(B, axis=1), 1e-300)
