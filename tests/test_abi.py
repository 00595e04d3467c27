"""The C-ABI libraries load without a GPU and export every function the
public headers declare (no compute calls)."""
import ctypes
import os
import re
import subprocess

import pytest

from cumf import _paths

INC = _paths.INCLUDE_DIR


def _declared(header):
    txt = open(os.path.join(INC, header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    txt = re.sub(r"//[^\n]*", "", txt)
    names = set(re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", txt))
    skip = {"if", "defined", "sizeof", "static_assert", "offsetof", "CUMF_JNI_SLOT", "return",
            "JNICALL", "Obj", "functions"}
    return {n for n in names if n not in skip}


def _exports(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True, check=True)
    return {line.split()[-1] for line in out.stdout.splitlines() if line.strip()}


def test_gpuals_exports_every_declared_entry_point():
    exp = _exports(_paths.ALS_LIB)
    want = (_declared("als.h") | {"get_gpu", "cumf_malloc_best", "cumf_free_best"}
            | {"Java_org_apache_spark_ml_recommendation_CuMFJNIInterface_doALSWithCSR",
               "Java_org_apache_spark_ml_recommendation_CuMFJNIInterface_testjni",
               "cumf_jni_flatten_factors", "cumf_jni_rows_to_array"})
    # C++-linkage symbols of the reference: allocation contract
    # (utilities.h:36-41) and the binary CSR loader (als/src/cuda/als.cu:914)
    want -= {"loadCSRSparseMatrixBin"}
    want |= {"_Z10mallocBestPPvm", "_Z8freeBestPv", "_Z22loadCSRSparseMatrixBinPKcS0_S0_PfPiS2_il"}
    missing = sorted(want - exp)
    assert not missing, missing
    assert "doALSWithCSR" in exp


def test_declared_names_cover_the_headers():
    names = _declared("als.h")
    assert {"doALSWithCSR", "cumf_create", "cumf_iterate", "cumf_heldout_rmse"} <= names


def test_data_lib_exports():
    exp = _exports(_paths.DATA_LIB)
    assert _declared("cumf_data.h") <= exp


def test_cpu_lib_exports():
    exp = _exports(_paths.CPU_LIB)
    assert _declared("cumf_cpu.h") <= exp


def test_libraries_load_without_gpu():
    ctypes.CDLL(_paths.ALS_LIB)
    ctypes.CDLL(_paths.DATA_LIB)
    ctypes.CDLL(_paths.CPU_LIB)


def test_product_does_not_link_the_oracle():
    out = subprocess.run(["ldd", _paths.ALS_LIB], capture_output=True, text=True).stdout
    assert "oracle" not in out
    exp = _exports(_paths.ALS_LIB)
    assert not any(s.startswith("oracle_") for s in exp)


def test_no_device_no_crash():
    """Without a visible GPU the context API returns an error, never exits."""
    from cumf import als as G
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is visible")
    with pytest.raises(G.CumfError):
        G.ALS(16, 0.1)
    assert G.lib().get_gpu() == -1
