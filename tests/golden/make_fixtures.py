"""Writes tests/golden/als_fixtures.npz: tiny ALS cases with expected outputs.

The reference ships no tests, golden vectors or fixtures for this path, so
the expected values here come from an independent dense numpy/scipy
computation of the definitions the reference implements (not from the
FP64 oracle under oracle/, which these fixtures pin):
  A_u = Y_u^T Y_u + fl32(lambda * n_u) I        als/src/cuda/als.cu:47-51, 905-909
  b_u = Y_u^T r_u                                als/src/cuda/als.cu:1039-1054
  x_u = A_u^{-1} b_u (x_u = 0 when n_u = 0)       als/src/cuda/als.cu:935-997
  compat: A(10I+4, 10J) := G(10I, 10J), I <= J   als/src/cuda/als.cu:275, 718
and of this library's iteration/RMSE definitions (X sweep, then Theta sweep;
RMSE over a held-out COO set, no clipping).

Run:  python tests/golden/make_fixtures.py   (numpy + scipy, seconds)
"""
import os

import numpy as np
import scipy.linalg as sl

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "als_fixtures.npz")


def make_csr(rng, m, n, per_row, empty_rows):
    rows, cols, vals = [], [], []
    for u in range(m):
        if u in empty_rows:
            continue
        k = int(rng.integers(1, per_row + 1))
        c = np.sort(rng.choice(n, size=min(k, n), replace=False))
        rows += [u] * len(c)
        cols += list(c)
        vals += list(rng.integers(1, 6, size=len(c)).astype(np.float32))
    rows = np.array(rows)
    rp = np.zeros(m + 1, np.int32)
    np.add.at(rp, rows + 1, 1)
    rp = np.cumsum(rp).astype(np.int32)
    return rp, np.array(cols, np.int32), np.array(vals, np.float32)


def dense_gram(rp, ci, cv, theta, lam, u):
    s, e = rp[u], rp[u + 1]
    Y = theta[ci[s:e]].astype(np.float64)
    r = cv[s:e].astype(np.float64)
    G = Y.T @ Y
    reg = float(np.float32(np.float32(lam) * np.float32(e - s)))
    return G, Y.T @ r, reg


def slip(G, f):
    A = G.copy()
    for I in range(f // 10):
        for J in range(I, f // 10):
            A[10 * I + 4, 10 * J] = G[10 * I, 10 * J]
    return A


def sweep(rp, ci, cv, theta, lam, compat=False):
    m, f = rp.shape[0] - 1, theta.shape[1]
    X = np.zeros((m, f))
    for u in range(m):
        if rp[u + 1] == rp[u]:
            continue
        G, b, reg = dense_gram(rp, ci, cv, theta, lam, u)
        if compat:
            X[u] = np.linalg.solve(slip(G, f) + reg * np.eye(f), b)
        else:
            X[u] = sl.cho_solve(sl.cho_factor(G + reg * np.eye(f)), b)
    return X


def transpose(rp, ci, cv, n):
    m = rp.shape[0] - 1
    rows = np.repeat(np.arange(m), np.diff(rp))
    order = np.lexsort((rows, ci))
    cp = np.zeros(n + 1, np.int32)
    np.add.at(cp, ci + 1, 1)
    return np.cumsum(cp).astype(np.int32), rows[order].astype(np.int32), cv[order]


def main():
    rng = np.random.default_rng(20240601)
    m, n = 40, 24
    rp, ci, cv = make_csr(rng, m, n, per_row=12, empty_rows={5, 31})
    cp, ri, cvt = transpose(rp, ci, cv, n)
    tu = rng.integers(0, m, 30).astype(np.int32)
    tv = rng.integers(0, n, 30).astype(np.int32)
    tr = rng.integers(1, 6, 30).astype(np.float32)
    lam = 0.058
    out = dict(rp=rp, ci=ci, cv=cv, cp=cp, ri=ri, cvt=cvt, tu=tu, tv=tv, tr=tr,
               lam=np.float64(lam), m=np.int64(m), n=np.int64(n))
    for f in (1, 4, 10, 32, 100):
        theta = rng.standard_normal((n, f)).astype(np.float32) / np.float32(np.sqrt(f))
        out[f"theta_f{f}"] = theta
        rows = [0, 1, 2] if f == 100 else list(range(m))
        A = np.zeros((len(rows), f, f))
        b = np.zeros((len(rows), f))
        for k, u in enumerate(rows):
            G, bu, reg = dense_gram(rp, ci, cv, theta, lam, u)
            A[k] = G + reg * np.eye(f)
            b[k] = bu
        out[f"A_f{f}"], out[f"b_f{f}"], out[f"Arows_f{f}"] = A, b, np.array(rows, np.int32)
        X1 = sweep(rp, ci, cv, theta, lam)
        out[f"X1_f{f}"] = X1
        # two full iterations in FP64 (X sweep, then Theta sweep), RMSE after each
        T = theta.astype(np.float64)
        hist = []
        for _ in range(2):
            X = sweep(rp, ci, cv, T, lam)
            T = sweep(cp, ri, cvt, X, lam)
            pred = np.einsum("ij,ij->i", X[tu], T[tv])
            hist.append(np.sqrt(np.mean((tr.astype(np.float64) - pred) ** 2)))
        out[f"X2_f{f}"], out[f"T2_f{f}"], out[f"rmse2_f{f}"] = X, T, np.array(hist)
        if f % 10 == 0:
            out[f"X1compat_f{f}"] = sweep(rp, ci, cv, theta, lam, compat=True)
            G, _, reg = dense_gram(rp, ci, cv, theta, lam, 0)
            out[f"A0compat_f{f}"] = slip(G, f) + reg * np.eye(f)
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
