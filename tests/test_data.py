"""Synthetic workload generator (libcumf_data.so): shape, law and
determinism properties."""
import numpy as np

import cumf
from tests import _data


def _check_csr(rp, ci, m, n):
    assert rp[0] == 0 and np.all(np.diff(rp) >= 0)
    for u in range(m):
        c = ci[rp[u]:rp[u + 1]]
        assert np.all(np.diff(c) > 0)  # sorted, distinct
    assert ci.min() >= 0 and ci.max() < n


def test_small_zipf_shape_and_split():
    ds = _data.small_zipf()
    assert (ds.m, ds.n) == (10000, 2000)
    assert ds.nnz + ds.test_u.shape[0] == 500000
    assert ds.test_u.shape[0] == 50000
    _check_csr(ds.csr_row_ptr, ds.csr_col, ds.m, ds.n)
    _check_csr(ds.csc_col_ptr, ds.csc_row, ds.n, ds.m)
    # every row and column keeps >= 1 training rating
    assert np.diff(ds.csr_row_ptr).min() >= 1
    assert np.diff(ds.csc_col_ptr).min() >= 1
    # held-out pairs are not training pairs
    train = set(zip(np.repeat(np.arange(ds.m), np.diff(ds.csr_row_ptr)).tolist(), ds.csr_col.tolist()))
    assert not any((u, v) in train for u, v in zip(ds.test_u[:5000].tolist(), ds.test_v[:5000].tolist()))
    # ratings are integers 1..5; literal latent model -> mostly ones
    assert set(np.unique(ds.csr_val).tolist()) <= {1.0, 2.0, 3.0, 4.0, 5.0}
    assert ds.info["hist"][0] / ds.nnz > 0.85


def test_csc_is_transpose_of_csr():
    ds = _data.netflix_like_small()
    rows = np.repeat(np.arange(ds.m), np.diff(ds.csr_row_ptr))
    cols = np.repeat(np.arange(ds.n), np.diff(ds.csc_col_ptr))
    a = sorted(zip(rows.tolist(), ds.csr_col.tolist(), ds.csr_val.tolist()))
    b = sorted(zip(ds.csc_row.tolist(), cols.tolist(), ds.csc_val.tolist()))
    assert a == b


def test_degree_law_is_power_law():
    ds = _data.small_zipf()
    deg = np.sort(np.diff(ds.csr_row_ptr))[::-1].astype(float)
    # Zipf-Mandelbrot with s = 1.1: log-log slope of the tail close to -1.1
    i = np.arange(deg.shape[0]) + ds.info["q_row"]
    sel = slice(100, 5000)
    slope = np.polyfit(np.log(i[sel]), np.log(deg[sel]), 1)[0]
    assert -1.3 < slope < -0.9


def test_deterministic_and_seeded():
    p = _data.params(300, 120, 6100, 100, row_cap=120, col_cap=200, seed=3)
    a = cumf.generate(p)
    b = cumf.generate(p)
    np.testing.assert_array_equal(a.csr_col, b.csr_col)
    np.testing.assert_array_equal(a.csr_val, b.csr_val)
    np.testing.assert_array_equal(a.test_u, b.test_u)
    p.seed = 4
    c = cumf.generate(p)
    assert not np.array_equal(a.csr_col, c.csr_col)


def test_presets_exist():
    p = cumf.preset("netflix_480k_17770_99m")
    assert (p.m, p.n, p.nnz_total - p.n_holdout) == (480189, 17770, 99072112)
    p = cumf.preset("zipf_4m_500k_1b")
    assert (p.m, p.n, p.nnz_total - p.n_holdout) == (4000000, 500000, 1000000000)


def test_init_factors_unit_norm():
    T = cumf.init_factors(50, 100, 1)
    np.testing.assert_allclose(np.linalg.norm(T, axis=1), 1.0, rtol=1e-5)
    np.testing.assert_array_equal(T, cumf.init_factors(50, 100, 1))
