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


def test_runals_usage_and_cpu_mode():
    exe = os.path.join(_paths.PKG_DIR, "runals")
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr
    r = subprocess.run([exe, "cpu", "10", "1", "0.1", "zipf_10k_2k_500k"], capture_output=True, text=True)
    assert r.returncode == 2 and "Spark" in r.stderr
