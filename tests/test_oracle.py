"""Pins the FP64 CPU oracle (oracle/) before it is trusted as the parity
oracle.  Each check is independent of the oracle's own code: committed
golden vectors from a dense numpy/scipy computation, scipy solvers,
closed forms, invariants and special cases, chosen so that a dropped term
(ridge, RHS), a wrong index (transposed Gram, wrong factor row) or a wrong
sign fails at least one of them."""
import os

import numpy as np
import pytest
import scipy.linalg as sl

from oracle import oracle as O

FIX = np.load(os.path.join(os.path.dirname(__file__), "golden", "als_fixtures.npz"))
LAM = float(FIX["lam"])
M, N = int(FIX["m"]), int(FIX["n"])
FS = (1, 4, 10, 32, 100)


@pytest.mark.parametrize("f", FS)
def test_row_system_matches_golden(f):
    theta = FIX[f"theta_f{f}"]
    for k, u in enumerate(FIX[f"Arows_f{f}"]):
        A, b = O.row_system(FIX["rp"], FIX["ci"], FIX["cv"], theta, f, LAM, u)
        np.testing.assert_allclose(A, FIX[f"A_f{f}"][k], rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(b, FIX[f"b_f{f}"][k], rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("f", FS)
def test_half_sweep_matches_golden(f):
    X, st = O.half_sweep(FIX["rp"], FIX["ci"], FIX["cv"], FIX[f"theta_f{f}"].astype(np.float64), LAM)
    assert st.max() == 0
    np.testing.assert_allclose(X, FIX[f"X1_f{f}"], rtol=1e-9, atol=1e-11)
    assert np.all(X[[5, 31]] == 0)  # rows without ratings


@pytest.mark.parametrize("f", FS)
def test_two_iterations_and_rmse_match_golden(f):
    class DS:
        pass
    ds = DS()
    ds.m, ds.n = M, N
    ds.csr_row_ptr, ds.csr_col, ds.csr_val = FIX["rp"], FIX["ci"], FIX["cv"]
    ds.csc_col_ptr, ds.csc_row, ds.csc_val = FIX["cp"], FIX["ri"], FIX["cvt"]
    ds.test_u, ds.test_v, ds.test_r = FIX["tu"], FIX["tv"], FIX["tr"]
    X, T, hist = O.als_train(ds, f, LAM, 2, FIX[f"theta_f{f}"])
    np.testing.assert_allclose(X, FIX[f"X2_f{f}"], rtol=1e-7, atol=1e-9)
    np.testing.assert_allclose(T, FIX[f"T2_f{f}"], rtol=1e-7, atol=1e-9)
    np.testing.assert_allclose(hist, FIX[f"rmse2_f{f}"], rtol=1e-10)


@pytest.mark.parametrize("f", [10, 100])
def test_compat_slip_matches_golden(f):
    theta = FIX[f"theta_f{f}"]
    A, _ = O.row_system(FIX["rp"], FIX["ci"], FIX["cv"], theta, f, LAM, 0, compat=True)
    np.testing.assert_allclose(A, FIX[f"A0compat_f{f}"], rtol=1e-12, atol=1e-12)
    X, st = O.half_sweep(FIX["rp"], FIX["ci"], FIX["cv"], theta.astype(np.float64), LAM, compat=True)
    assert st.max() == 0
    np.testing.assert_allclose(X, FIX[f"X1compat_f{f}"], rtol=1e-8, atol=1e-10)


def test_compat_touches_exactly_55_entries_at_f100():
    """The reference's slip (als.cu:718) overwrites one entry per upper 10x10
    tile (55 tiles at f=100): the entry (10I+4, 10J) holds G(10I, 10J), the
    unregularised Gram value; every other entry is the clean one."""
    f = 100
    theta = FIX["theta_f100"]
    rp, ci, cv = FIX["rp"], FIX["ci"], FIX["cv"]
    A, _ = O.row_system(rp, ci, cv, theta, f, LAM, 0)
    Ac, _ = O.row_system(rp, ci, cv, theta, f, LAM, 0, compat=True)
    ridge = float(np.float32(np.float32(LAM) * np.float32(rp[1] - rp[0])))
    G = A - ridge * np.eye(f)
    expect = {(10 * I + 4, 10 * J) for I in range(10) for J in range(I, 10)}
    assert len(expect) == 55
    mask = np.zeros((f, f), bool)
    for (i, j) in expect:
        mask[i, j] = True
        assert abs(Ac[i, j] - G[i - 4, j]) <= 1e-12 * max(1.0, abs(G[i - 4, j]))
    np.testing.assert_array_equal(Ac[~mask], A[~mask])
    assert int(np.count_nonzero(Ac != A)) == 55
    assert not np.array_equal(Ac, Ac.T)


def test_compat_rejects_f_not_multiple_of_10():
    with pytest.raises(ValueError):
        O.row_system(FIX["rp"], FIX["ci"], FIX["cv"], FIX["theta_f32"], 32, LAM, 0, compat=True)


@pytest.mark.parametrize("f", [1, 4, 10, 32])
def test_row_system_invariants(f):
    """Symmetric, SPD, diag >= ridge, trace = sum |theta_v|^2 + f ridge."""
    theta = FIX[f"theta_f{f}"]
    rp, ci = FIX["rp"], FIX["ci"]
    for u in range(M):
        A, b = O.row_system(rp, ci, FIX["cv"], theta, f, LAM, u)
        s, e = rp[u], rp[u + 1]
        ridge = float(np.float32(np.float32(LAM) * np.float32(e - s)))
        assert np.array_equal(A, A.T)
        if e > s:
            assert np.linalg.eigvalsh(A).min() > 0
        assert np.all(np.diag(A) >= ridge)
        tr = float(np.sum(theta[ci[s:e]].astype(np.float64) ** 2)) + f * ridge
        assert abs(np.trace(A) - tr) <= 1e-12 * max(tr, 1.0)


def test_f1_reduces_to_scalars():
    theta = FIX["theta_f1"]
    rp, ci, cv = FIX["rp"], FIX["ci"], FIX["cv"]
    for u in (0, 3, 7):
        s, e = rp[u], rp[u + 1]
        t = theta[ci[s:e], 0].astype(np.float64)
        A, b = O.row_system(rp, ci, cv, theta, 1, LAM, u)
        ridge = float(np.float32(np.float32(LAM) * np.float32(e - s)))
        assert abs(A[0, 0] - (np.sum(t * t) + ridge)) < 1e-12
        assert abs(b[0] - np.sum(cv[s:e] * t)) < 1e-12


def test_unit_vector_factors_give_count_matrix():
    """theta_v = e_{v mod f}: A = diag(counts) + ridge I, b = per-slot rating sums."""
    f = 4
    theta = np.zeros((N, f), np.float32)
    theta[np.arange(N), np.arange(N) % f] = 1.0
    rp, ci, cv = FIX["rp"], FIX["ci"], FIX["cv"]
    for u in (0, 1, 2):
        s, e = rp[u], rp[u + 1]
        A, b = O.row_system(rp, ci, cv, theta, f, LAM, u)
        cnt = np.bincount(ci[s:e] % f, minlength=f)
        sums = np.bincount(ci[s:e] % f, weights=cv[s:e], minlength=f)
        ridge = float(np.float32(np.float32(LAM) * np.float32(e - s)))
        np.testing.assert_allclose(A, np.diag(cnt + ridge), atol=1e-14)
        np.testing.assert_allclose(b, sums, atol=1e-12)


def test_cholesky_and_lu_against_scipy():
    rng = np.random.default_rng(1)
    for f in (1, 3, 17, 100):
        B = rng.standard_normal((f + 5, f))
        A = B.T @ B + 0.1 * np.eye(f)
        b = rng.standard_normal(f)
        np.testing.assert_allclose(O.chol_solve(A, b), sl.cho_solve(sl.cho_factor(A), b), rtol=1e-9, atol=1e-12)
        N_ = rng.standard_normal((f, f)) + (2 * f) * np.eye(f)  # non-symmetric, no pivot needed
        np.testing.assert_allclose(O.lu_solve(N_, b), sl.lu_solve(sl.lu_factor(N_), b), rtol=1e-10, atol=1e-12)
    with pytest.raises(np.linalg.LinAlgError):
        O.chol_solve(-np.eye(3), np.ones(3))


@pytest.mark.parametrize("f", [4, 32])
def test_half_sweep_is_the_regularised_least_squares_solution(f):
    """x_u minimises |Y x - r|^2 + ridge |x|^2: the stacked lstsq solution,
    and the gradient Y^T(Yx - r) + ridge x vanishes."""
    theta = FIX[f"theta_f{f}"].astype(np.float64)
    rp, ci, cv = FIX["rp"], FIX["ci"], FIX["cv"]
    X, _ = O.half_sweep(rp, ci, cv, theta, LAM)
    for u in range(M):
        s, e = rp[u], rp[u + 1]
        if e == s:
            continue
        Y = theta[ci[s:e]]
        r = cv[s:e].astype(np.float64)
        ridge = float(np.float32(np.float32(LAM) * np.float32(e - s)))
        S = np.vstack([Y, np.sqrt(ridge) * np.eye(f)])
        xs = np.linalg.lstsq(S, np.concatenate([r, np.zeros(f)]), rcond=None)[0]
        np.testing.assert_allclose(X[u], xs, rtol=1e-8, atol=1e-10)
        g = Y.T @ (Y @ X[u] - r) + ridge * X[u]
        assert np.linalg.norm(g) <= 1e-9 * max(1.0, np.linalg.norm(Y.T @ r))


def test_rmse_closed_form():
    X = np.array([[1.0, 0.0], [0.0, 2.0]])
    T = np.array([[3.0, 1.0], [1.0, 1.0]])
    u = np.array([0, 1, 1], np.int32)
    v = np.array([0, 0, 1], np.int32)
    r = np.array([4.0, 1.0, 2.0], np.float32)
    # predictions 3, 2, 2 -> errors 1, -1, 0
    assert abs(O.rmse(u, v, r, X, T) - np.sqrt(2.0 / 3.0)) < 1e-15
    assert abs(O.rmse(u, v, r, X.astype(np.float32), T.astype(np.float32)) - np.sqrt(2.0 / 3.0)) < 1e-7


def test_objective_never_increases_across_exact_half_sweeps():
    f = 10
    rp, ci, cv = FIX["rp"], FIX["ci"], FIX["cv"]
    cp, ri, cvt = FIX["cp"], FIX["ri"], FIX["cvt"]
    T = FIX["theta_f10"].astype(np.float64)
    X = np.zeros((M, f))
    prev = np.inf
    for _ in range(4):
        X, _ = O.half_sweep(rp, ci, cv, T, LAM)
        obj = O.objective(rp, ci, cv, X, T, LAM)
        assert obj <= prev * (1 + 1e-12)
        prev = obj
        T, _ = O.half_sweep(cp, ri, cvt, X, LAM)
        obj = O.objective(rp, ci, cv, X, T, LAM)
        assert obj <= prev * (1 + 1e-12)
        prev = obj


def test_objective_closed_form():
    rp = np.array([0, 1, 2], np.int32)
    ci = np.array([0, 1], np.int32)
    cv = np.array([3.0, 1.0], np.float32)
    X = np.array([[1.0], [2.0]])
    T = np.array([[2.0], [1.0]])
    # loss: (3-2)^2 + (1-2)^2 = 2; reg: lam*(1*1 + 1*4 + 1*4 + 1*1) = 10 lam
    lam = 0.5
    assert abs(O.objective(rp, ci, cv, X, T, lam) - (2 + 10 * lam)) < 1e-12


@pytest.mark.parametrize("f", [10, 100])
def test_compat_system_reference_arithmetic(f):
    """The FP32 sequential-FMA system agrees with the FP64 clean system to
    FP32 rounding away from the slipped entries, carries the slip exactly,
    and solving it by LU reproduces the golden compat solution closely."""
    theta = FIX[f"theta_f{f}"]
    rp, ci, cv = FIX["rp"], FIX["ci"], FIX["cv"]
    for u in (0, 1, 2):
        M, b = O.compat_system(rp, ci, cv, theta, f, LAM, u)
        Ac, bc = O.row_system(rp, ci, cv, theta, f, LAM, u, compat=True)
        n_u = rp[u + 1] - rp[u]
        tol = 4 * n_u * np.finfo(np.float32).eps * (np.abs(Ac).max() + 1)
        assert np.abs(M - Ac).max() <= tol
        assert np.abs(b - bc).max() <= 4 * n_u * np.finfo(np.float32).eps * (np.abs(bc).max() + 5)
        ridge = float(np.float32(np.float32(n_u) * np.float32(LAM)))
        for I in range(f // 10):
            for J in range(I, f // 10):
                src = M[10 * I, 10 * J] - (ridge if I == J else 0.0)
                assert abs(M[10 * I + 4, 10 * J] - src) <= (1e-12 * abs(src) if I == J else 0.0)
        # sequential fmaf: entries of symmetric positions agree bit for bit
        G = M - np.diag(np.diag(M))
        mask = np.ones((f, f), bool)
        for I in range(f // 10):
            for J in range(I, f // 10):
                mask[10 * I + 4, 10 * J] = mask[10 * J, 10 * I + 4] = False
        np.testing.assert_array_equal(G[mask], G.T[mask])
    x = O.lu_solve(*O.compat_system(rp, ci, cv, theta, f, LAM, 0))
    np.testing.assert_allclose(x, FIX[f"X1compat_f{f}"][0], rtol=1e-4, atol=1e-5)
