"""FP64 numpy restatement of the warm-start MLP (forward, backward, Adam).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the oracle of the HIP
MLP kernels, never by the product path.

The network maps the per-row features of one half-sweep,
    [log n_u, diag(A_u) / n_u, b_u / n_u]            (2f + 1 values),
to a predicted solution x_u (f values):
    H0 = bf16(features)                    zero-padded to K0 = round_up(2f+1, 32)
    Z1 = H0 W1^T + c1,  H1 = bf16(gelu(Z1))           (hidden = 256)
    Z2 = H1 W2^T + c2,  H2 = bf16(gelu(Z2))
    Y  = H2 W3^T + c3                                 (N3 = round_up(f, 32) rows of W3)
with BF16 copies of the weights as GEMM operands and FP32 accumulation;
gelu(z) = z Phi(z) (erf form).  Loss: mean over the batch and the f real
outputs of (Y - x*)^2.  Backward (the rounding points the kernels use):
    dY  = bf16(2 (Y - x*) / (B f))   (0 on padded outputs)
    dW3 = dY^T H2,  dc3 = colsum(dY)
    dZ2 = bf16((dY W3) * gelu'(Z2)),  dW2 = dZ2^T H1, ...
    dZ1 = bf16((dZ2 W2) * gelu'(Z1)), dW1 = dZ1^T H0
Adam (Kingma & Ba 2015, Algorithm 1) on FP32 master weights.

This is the library's own network (the reference has no learned part); the
functions are pinned by central finite differences of the unrounded FP64
loss in tests/test_mlp_oracle.py.
"""
import math

import numpy as np

HIDDEN = 256


def dims(f):
    k0 = (2 * f + 1 + 31) // 32 * 32
    n3 = (f + 31) // 32 * 32
    return k0, HIDDEN, HIDDEN, n3


def bf16(x):
    """Round FP32 values to BF16 (nearest, ties to even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    nan = np.isnan(a)
    out = r.view(np.float32).astype(np.float64)
    out[nan] = np.nan
    return out


def _erf(z):
    return np.vectorize(math.erf, otypes=[np.float64])(z)


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + _erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def init_params(f, seed):
    """Uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights over the real inputs,
    zero biases; padded rows/columns zero.  Same generator as the library
    (splitmix64 stream, see mlp.cpp)."""
    k0, h1, h2, n3 = dims(f)
    shapes = [(h1, k0, 2 * f + 1, h1), (h2, h1, h1, h2), (n3, h2, h2, f)]
    params = []
    state = np.uint64(seed)
    for (n, k, kreal, nreal) in shapes:
        W = np.zeros((n, k), np.float32)
        bound = 1.0 / math.sqrt(kreal)
        vals, state = _splitmix_uniform(state, nreal * kreal)
        W[:nreal, :kreal] = ((2.0 * vals - 1.0) * bound).reshape(nreal, kreal)
        params.append(W)
        params.append(np.zeros(n, np.float32))
    return params


def _splitmix_uniform(state, count):
    """count uniforms in [0, 1) from a splitmix64 stream (top 24 bits)."""
    out = np.empty(count, np.float64)
    s = int(state)
    M = (1 << 64) - 1
    for i in range(count):
        s = (s + 0x9E3779B97F4A7C15) & M
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        z = z ^ (z >> 31)
        out[i] = (z >> 40) / float(1 << 24)
    return out.astype(np.float32), np.uint64(s)


def forward(params, X, emulate=True):
    """X: (B, K0) features (already padded).  Returns (Y, cache)."""
    rnd = bf16 if emulate else (lambda v: np.asarray(v, np.float64))
    W1, c1, W2, c2, W3, c3 = params
    H0 = rnd(X)
    Z1 = H0 @ rnd(W1).T + np.asarray(c1, np.float64)
    H1 = rnd(gelu(Z1))
    Z2 = H1 @ rnd(W2).T + np.asarray(c2, np.float64)
    H2 = rnd(gelu(Z2))
    Y = H2 @ rnd(W3).T + np.asarray(c3, np.float64)
    return Y, (H0, Z1, H1, Z2, H2)


def loss(Y, T, f):
    d = Y[:, :f] - T[:, :f]
    return float(np.mean(d * d))


def backward(params, cache, Y, T, f, emulate=True):
    """Gradients of the mean-squared loss: [dW1, dc1, dW2, dc2, dW3, dc3]."""
    rnd = bf16 if emulate else (lambda v: np.asarray(v, np.float64))
    W1, c1, W2, c2, W3, c3 = params
    H0, Z1, H1, Z2, H2 = cache
    B = Y.shape[0]
    dY = np.zeros_like(Y)
    dY[:, :f] = 2.0 * (Y[:, :f] - T[:, :f]) / (B * f)
    dY = rnd(dY)
    dW3 = dY.T @ H2
    dc3 = dY.sum(axis=0)
    dZ2 = rnd((dY @ rnd(W3)) * gelu_grad(Z2))
    dW2 = dZ2.T @ H1
    dc2 = dZ2.sum(axis=0)
    dZ1 = rnd((dZ2 @ rnd(W2)) * gelu_grad(Z1))
    dW1 = dZ1.T @ H0
    dc1 = dZ1.sum(axis=0)
    return [dW1, dc1, dW2, dc2, dW3, dc3]


def adam_step(params, grads, m, v, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (step t >= 1) in place on FP32 arrays."""
    for p, g, mm, vv in zip(params, grads, m, v):
        g = np.asarray(g, np.float32)
        mm[...] = b1 * mm + (1 - b1) * g
        vv[...] = b2 * vv + (1 - b2) * g * g
        mh = mm / (1 - b1 ** t)
        vh = vv / (1 - b2 ** t)
        p[...] = p - lr * mh / (np.sqrt(vh) + eps)


def features(row_ptr, col, val, theta, f, lam, rows):
    """FP64 features of the listed rows: [log n_u, diag(A_u)/n_u, b_u/n_u]
    with A_u = Y_u^T Y_u + lam n_u I, b_u = Y_u^T r_u; zero-padded to K0."""
    k0 = dims(f)[0]
    out = np.zeros((len(rows), k0), np.float64)
    th = np.asarray(theta, np.float64)
    for k, u in enumerate(rows):
        s, e = int(row_ptr[u]), int(row_ptr[u + 1])
        n = e - s
        if n == 0:
            continue
        Y = th[col[s:e], :f]
        r = np.asarray(val[s:e], np.float64)
        out[k, 0] = math.log(n)
        out[k, 1:1 + f] = (Y * Y).sum(axis=0) / n + lam
        out[k, 1 + f:1 + 2 * f] = (r @ Y) / n
    return out
