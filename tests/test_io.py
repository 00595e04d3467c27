"""The reference's raw binary CSR format (als/src/cuda/als.cu:914-933) and
the standalone driver's argument handling; CPU only."""
import os
import subprocess

import numpy as np
import pytest

from cumf import _paths
from cumf import als as G
from tests import _data


def test_csr_bin_round_trip(tmp_path):
    ds = _data.tiny()
    pre = str(tmp_path / "r")
    G.save_csr_bin(pre, ds.csr_row_ptr, ds.csr_col, ds.csr_val)
    assert os.path.getsize(pre + ".row.bin") == 4 * (ds.m + 1)
    assert os.path.getsize(pre + ".col.bin") == 4 * ds.nnz
    assert os.path.getsize(pre + ".data.bin") == 4 * ds.nnz
    rp, ci, cv = G.load_csr_bin(pre, ds.m, ds.nnz)
    np.testing.assert_array_equal(rp, ds.csr_row_ptr)
    np.testing.assert_array_equal(ci, ds.csr_col)
    np.testing.assert_array_equal(cv, ds.csr_val)
    # the files are exactly the raw arrays, no header
    np.testing.assert_array_equal(np.fromfile(pre + ".col.bin", np.int32), ds.csr_col)


def test_csr_bin_checks_every_file_and_size(tmp_path):
    ds = _data.tiny()
    pre = str(tmp_path / "r")
    G.save_csr_bin(pre, ds.csr_row_ptr, ds.csr_col, ds.csr_val)
    with pytest.raises(G.CumfError, match="size"):
        G.load_csr_bin(pre, ds.m, ds.nnz + 1)
    os.remove(pre + ".col.bin")  # the reference never checked this file
    with pytest.raises(G.CumfError, match="cannot open"):
        G.load_csr_bin(pre, ds.m, ds.nnz)


def test_reference_loader_symbol_is_exported():
    out = subprocess.run(["nm", "-D", "--defined-only", _paths.ALS_LIB], capture_output=True, text=True).stdout
    assert "_Z22loadCSRSparseMatrixBinPKcS0_S0_PfPiS2_il" in out


def test_runals_usage_and_option_checks():
    exe = os.path.join(_paths.PKG_DIR, "runals")
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr
    r = subprocess.run([exe, "cpu", "10", "1", "0.1", "zipf_10k_2k_500k", "--solver", "cg"], capture_output=True,
                       text=True)
    assert r.returncode == 2 and "gpu options" in r.stderr
    r = subprocess.run([exe, "cpu", "10", "1", "0.1", "zipf_10k_2k_500k", "--holdout", "0.1"], capture_output=True,
                       text=True)
    assert r.returncode == 2 and "--holdout" in r.stderr


def test_runals_cpu_mode_matches_fp64_oracle(tmp_path):
    """runALS.sh's cpu leg without Spark: the host half-sweeps from the same
    seeded start as the oracle's FP64 ALS, on ratings read from the
    reference's binary CSR files and split by --holdout."""
    import json

    from oracle import oracle as O
    ds = _data.netflix_like_small()
    pre = str(tmp_path / "r")
    G.save_csr_bin(pre, ds.csr_row_ptr, ds.csr_col, ds.csr_val)
    exe = os.path.join(_paths.PKG_DIR, "runals")
    f, lam, iters = 16, 0.058, 3
    r = subprocess.run([exe, "cpu", str(f), str(iters), str(lam), f"csr:{ds.m}:{ds.n}:{ds.nnz}:{pre}",
                        "--holdout", "0.05", "--seed", "0", "--threads", "4"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    summary = json.loads(r.stdout.strip().splitlines()[-1])
    assert summary["mode"] == "cpu" and summary["sweeps_with_failed_rows"] == 0
    trp, tci, tcv, hu, hv, hr = G.split_holdout(ds.csr_row_ptr, ds.csr_col, ds.csr_val, ds.n, 0.05, 0)
    assert summary["nnz"] == tci.shape[0]
    import cumf
    from cumf.data import Dataset
    order = np.argsort(tci, kind="stable")
    cp = np.zeros(ds.n + 1, np.int32)
    np.add.at(cp, tci + 1, 1)
    cp = np.cumsum(cp).astype(np.int32)
    rows = np.repeat(np.arange(ds.m, dtype=np.int32), np.diff(trp))
    split = Dataset(m=ds.m, n=ds.n, csr_row_ptr=trp, csr_col=tci, csr_val=tcv, csc_col_ptr=cp,
                    csc_row=rows[order], csc_val=tcv[order], test_u=hu, test_v=hv, test_r=hr, params={}, info={})
    _, _, hist = O.als_train(split, f, lam, iters, cumf.init_factors(ds.n, f, 0))
    assert abs(summary["heldout_rmse"] - hist[-1]) <= 1e-5, (summary["heldout_rmse"], hist[-1])


def _split_checks(rp, ci, cv, n, fraction, seed):
    trp, tci, tcv, hu, hv, hr = G.split_holdout(rp, ci, cv, n, fraction, seed)
    m = rp.shape[0] - 1
    assert trp[0] == 0 and trp[-1] == tci.shape[0]
    assert tci.shape[0] + hu.shape[0] == ci.shape[0]
    # train + held out is the original multiset of (u, v, r), order kept
    rows = np.repeat(np.arange(m), np.diff(rp))
    trows = np.repeat(np.arange(m), np.diff(trp))
    orig = set(zip(rows.tolist(), ci.tolist(), cv.tolist()))
    got = list(zip(trows.tolist(), tci.tolist(), tcv.tolist())) + list(zip(hu.tolist(), hv.tolist(), hr.tolist()))
    assert len(got) == len(orig) and set(got) == orig
    # every row and column with ratings keeps a training rating
    assert np.all(np.diff(trp)[np.diff(rp) > 0] > 0)
    ccount = np.bincount(ci, minlength=n)
    tcount = np.bincount(tci, minlength=n)
    assert np.all(tcount[ccount > 0] > 0)
    return trp, tci, tcv, hu, hv, hr


def test_split_holdout_properties():
    ds = _data.small_zipf()
    for fraction in (0.0, 0.1, 0.5, 1.0):
        trp, tci, _, hu, _, _ = _split_checks(ds.csr_row_ptr, ds.csr_col, ds.csr_val, ds.n, fraction, 7)
        if fraction == 0.0:
            assert hu.shape[0] == 0
        elif fraction == 1.0:
            # only the kept ratings remain: at most one per row plus one per column
            assert tci.shape[0] <= ds.m + ds.n
        else:
            got = hu.shape[0] / ds.nnz
            assert abs(got - fraction) < 0.01, got
    a = G.split_holdout(ds.csr_row_ptr, ds.csr_col, ds.csr_val, ds.n, 0.1, 7)
    b = G.split_holdout(ds.csr_row_ptr, ds.csr_col, ds.csr_val, ds.n, 0.1, 7)
    c = G.split_holdout(ds.csr_row_ptr, ds.csr_col, ds.csr_val, ds.n, 0.1, 8)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
    assert a[3].shape != c[3].shape or not np.array_equal(a[3], c[3])


def test_split_holdout_single_rating_rows_and_empty_rows():
    # rows 0 and 2 have one rating each, row 1 none; column 3 appears once
    rp = np.array([0, 1, 1, 2, 6], np.int32)
    ci = np.array([0, 1, 0, 1, 2, 3], np.int32)
    cv = np.array([1, 2, 3, 4, 5, 1], np.float32)
    trp, tci, tcv, hu, hv, hr = _split_checks(rp, ci, cv, 5, 1.0, 3)
    assert np.diff(trp).tolist()[:3] == [1, 0, 1]


def test_split_holdout_rejects_bad_input():
    rp = np.array([0, 2, 1], np.int32)
    with pytest.raises(G.CumfError):
        G.split_holdout(rp, np.array([0, 1], np.int32), np.ones(2, np.float32), 2, 0.1)
    with pytest.raises(G.CumfError):
        G.split_holdout(np.array([0, 1], np.int32), np.array([5], np.int32), np.ones(1, np.float32), 2, 0.1)
    with pytest.raises(G.CumfError):
        G.split_holdout(np.array([0, 1], np.int32), np.array([0], np.int32), np.ones(1, np.float32), 2, 1.5)
