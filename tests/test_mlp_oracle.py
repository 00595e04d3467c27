"""Pins the FP64 MLP oracle (oracle/mlp_oracle.py) on the CPU.

  * backward(): central finite differences of the unrounded FP64 loss on a
    small network (f=4: 9 -> 256 -> 256 -> 4 real units) agree with the
    analytic gradient of every parameter tensor to 1e-6 relative;
  * gelu / gelu_grad: closed forms at 0 and the large-|z| limits, and
    gelu_grad against a finite difference of gelu;
  * bf16(): exact on BF16-representable values, ties to even, relative
    error <= 2^-8 (half a unit of the 8-bit significand) elsewhere;
  * adam_step(): the first step moves each weight by lr * sign(g)
    (m_hat / sqrt(v_hat) = g / |g| at t = 1);
  * features(): a row with one rating of unit factor gives diag = 1 + lam.
"""
import math

import numpy as np

from oracle import mlp_oracle as M


def _net(f=4, seed=3, B=5):
    rng = np.random.default_rng(seed)
    params = M.init_params(f, seed)
    # non-zero biases so their gradients are exercised
    for i in (1, 3, 5):
        params[i][:] = rng.normal(0, 0.1, params[i].shape).astype(np.float32)
    k0 = M.dims(f)[0]
    X = np.zeros((B, k0))
    X[:, :2 * f + 1] = rng.normal(0, 1, (B, 2 * f + 1))
    T = rng.normal(0, 1, (B, M.dims(f)[3]))
    return f, params, X, T


def test_backward_matches_finite_differences():
    f, params, X, T = _net()
    P = [np.asarray(p, np.float64) for p in params]
    Y, cache = M.forward(P, X, emulate=False)
    grads = M.backward(P, cache, Y, T, f, emulate=False)
    rng = np.random.default_rng(0)
    h = 1e-6
    for pi, (p, g) in enumerate(zip(P, grads)):
        flat = p.reshape(-1)
        # only entries that influence the loss (real rows/columns)
        idx = rng.choice(flat.size, size=min(25, flat.size), replace=False)
        for j in idx:
            old = flat[j]
            flat[j] = old + h
            lp = M.loss(M.forward(P, X, emulate=False)[0], T, f)
            flat[j] = old - h
            lm = M.loss(M.forward(P, X, emulate=False)[0], T, f)
            flat[j] = old
            fd = (lp - lm) / (2 * h)
            an = g.reshape(-1)[j]
            assert abs(fd - an) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (pi, j, fd, an)


def test_gelu_closed_forms():
    z = np.array([0.0, 8.0, -8.0, 1.0])
    g = M.gelu(z)
    assert g[0] == 0.0 and abs(g[1] - 8.0) < 1e-12 and abs(g[2]) < 1e-12
    assert abs(g[3] - 0.5 * (1 + math.erf(1 / math.sqrt(2)))) < 1e-15
    d = M.gelu_grad(np.array([0.0, 8.0, -8.0]))
    assert abs(d[0] - 0.5) < 1e-15 and abs(d[1] - 1.0) < 1e-12 and abs(d[2]) < 1e-12
    zz = np.linspace(-4, 4, 41)
    fd = (M.gelu(zz + 1e-6) - M.gelu(zz - 1e-6)) / 2e-6
    assert np.abs(fd - M.gelu_grad(zz)).max() < 1e-8


def test_bf16_rounding():
    exact = np.array([1.0, -2.5, 0.15625, 2.0 ** 127, 3 * 2.0 ** -130], np.float32)
    assert np.array_equal(M.bf16(exact), exact.astype(np.float64))
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties to even -> 1
    assert M.bf16(np.float32(1 + 2 ** -8)) == 1.0
    # 1 + 3 * 2^-8 is halfway between 1 + 2^-7 and 1 + 2^-6: ties to even -> 1 + 2^-6
    assert M.bf16(np.float32(1 + 3 * 2 ** -8)) == 1 + 2 ** -6
    x = np.random.default_rng(1).normal(0, 10, 1000).astype(np.float32)
    rel = np.abs(M.bf16(x) - x) / np.abs(x)
    assert rel.max() <= 2 ** -8


def test_adam_first_step_is_signed_lr():
    p = [np.ones(6, np.float32)]
    g = [np.array([3.0, -2.0, 0.5, -0.25, 1e-3, -7.0], np.float32)]
    m = [np.zeros(6, np.float32)]
    v = [np.zeros(6, np.float32)]
    M.adam_step(p, g, m, v, 1, lr=1e-2)
    assert np.allclose(p[0], 1.0 - 1e-2 * np.sign(g[0]), atol=1e-6)


def test_features_single_unit_rating():
    f, lam = 3, 0.25
    theta = np.array([[1.0, 0.0, 0.0], [0.0, 2.0, 0.0]], np.float32)
    row_ptr = np.array([0, 1, 3], np.int32)
    col = np.array([0, 0, 1], np.int32)
    val = np.array([4.0, 1.0, 3.0], np.float32)
    F = M.features(row_ptr, col, val, theta, f, lam, [0, 1])
    assert F[0, 0] == 0.0
    assert np.allclose(F[0, 1:4], [1 + lam, lam, lam]) and np.allclose(F[0, 4:7], [4, 0, 0])
    assert abs(F[1, 0] - math.log(2)) < 1e-15
    assert np.allclose(F[1, 1:4], [0.5 + lam, 2.0 + lam, lam]) and np.allclose(F[1, 4:7], [0.5, 3.0, 0])
