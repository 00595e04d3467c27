"""Writes tests/golden/netflix_f100_fp64.npz: the FP64 ALS trajectory of the
Netflix-shaped workload (generator preset netflix_480k_17770_99m:
480,189 x 17,770, 99,072,112 training ratings, 990,721 held out), f=100,
lambda=0.058, 10 iterations (X sweep over the CSR, then Theta sweep over the
CSC) from init_factors(17770, 100, seed 0).

Stored: CRC-32 checksums of every generated array (so a GPU test knows it
regenerated the same ratings), the held-out RMSE after each iteration, and
the factor rows of a fixed sample of users and items after iterations 2 and
10 (FP64, stored as float64).

The expected values come from the FP64 CPU oracle (oracle/), whose
functions the tiny fixtures of make_fixtures.py and the closed-form tests
of tests/test_oracle.py pin; this file extends it to the full-size workload
that a CPU run can still afford (about ten minutes on eight cores).

Run:  python tests/golden/make_workload_fixtures.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
for p in (os.path.join(REPO, "cuda-mllib_amd"), REPO):
    if p not in sys.path:
        sys.path.insert(0, p)

import cumf  # noqa: E402
from oracle import oracle as O  # noqa: E402
from tests._checks import DATASET_ARRAYS, dataset_crc  # noqa: E402

OUT = os.path.join(HERE, "netflix_f100_fp64.npz")


def sample_rows(row_ptr, k, seed):
    """k rows spread over the degree range: the heaviest, the lightest and a
    seeded uniform draw from the rest."""
    deg = np.diff(row_ptr)
    order = np.argsort(-deg, kind="stable")
    pick = list(order[:k // 8]) + list(order[-(k // 8):])
    rest = np.setdiff1d(np.arange(deg.shape[0]), pick)
    pick += list(np.random.default_rng(seed).choice(rest, k - len(pick), replace=False))
    return np.sort(np.array(pick, np.int64))


def main():
    f, lam, iters = 100, 0.058, 10
    t0 = time.time()
    ds = cumf.generate(cumf.preset("netflix_480k_17770_99m"))
    print(f"generated in {time.time() - t0:.1f}s: m={ds.m} n={ds.n} nnz={ds.nnz}", flush=True)
    theta0 = cumf.init_factors(ds.n, f, 0)
    users = sample_rows(ds.csr_row_ptr, 256, 1)
    items = sample_rows(ds.csc_col_ptr, 64, 2)
    t0 = time.time()
    X2, T2, h2 = O.als_train(ds, f, lam, 2, theta0)
    print(f"2 iterations in {time.time() - t0:.1f}s: rmse {h2}", flush=True)
    t0 = time.time()
    X10, T10, hist = O.als_train(ds, f, lam, iters, theta0)
    print(f"{iters} iterations in {time.time() - t0:.1f}s: rmse {hist}", flush=True)
    assert np.array_equal(hist[:2], h2)
    np.savez_compressed(
        OUT, preset=np.array("netflix_480k_17770_99m"), f=np.int64(f), lam=np.float64(lam),
        iters=np.int64(iters), init_seed=np.int64(0), crc=dataset_crc(ds), crc_arrays=np.array(DATASET_ARRAYS),
        rmse=hist, users=users, items=items, X2=X2[users], T2=T2[items], X10=X10[users], T10=T10[items])
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
