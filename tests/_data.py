"""Small seeded datasets shared by the tests (cached per process)."""
import functools

import cumf
from cumf.data import GenParams


def params(m, n, nnz_total, n_holdout, row_s=1.0, col_s=0.8, row_cap=None, col_cap=None, rank=20,
           latent_var=0.2236, noise_std=0.7, mu=3.6, seed=0):
    return GenParams(m, n, nnz_total, n_holdout, row_s, col_s, row_cap or n, col_cap or m, rank,
                     latent_var, noise_std, mu, 1, seed)


@functools.lru_cache(maxsize=None)
def small_zipf():
    """m=10,000, n=2,000, 500,000 ratings (10% held out), Zipf(1.1)."""
    return cumf.generate(cumf.preset("zipf_10k_2k_500k"))


@functools.lru_cache(maxsize=None)
def netflix_like_small(seed=0):
    """m=4,000, n=1,500, 400,000 ratings (1% held out); rows Zipf(1.0) up
    to 1,500 ratings, columns Zipf(0.8): heavy rows and columns."""
    return cumf.generate(params(4000, 1500, 404000, 4000, row_cap=1500, col_cap=2000, seed=seed))


@functools.lru_cache(maxsize=None)
def tiny(seed=0):
    """m=300, n=120, 6,000 ratings."""
    return cumf.generate(params(300, 120, 6100, 100, row_cap=120, col_cap=200, seed=seed))
