"""Spark JNI bridge marshalling, driven through a fake JNIEnv (no JVM, no
GPU): factor flattening order (als/src/CuMFJNIInterface.cpp:60-76), the
float[m][f] result (:111-118) and error paths that raise Java exceptions."""
import ctypes as C
import os

import numpy as np
import pytest

from cumf import als as G

FJ = C.CDLL(os.path.join(os.path.dirname(__file__), "jni", "libfakejni.so"))
P = C.c_void_p
for name, args, res in [
    ("fj_create", [], P), ("fj_destroy", [P], None), ("fj_env", [P], P),
    ("fj_new_floats", [P, P, C.c_int], P), ("fj_new_ints", [P, P, C.c_int], P),
    ("fj_new_objects", [P, C.c_int], P), ("fj_set", [P, C.c_int, P], None),
    ("fj_get", [P, C.c_int], P), ("fj_len", [P], C.c_int), ("fj_floats", [P, P, C.c_int], C.c_int),
    ("fj_exception", [P, C.c_char_p, C.c_int, C.c_char_p, C.c_int], C.c_int),
    ("fj_clear_exception", [P], None), ("fj_critical_depth", [P], C.c_int),
    ("fj_fail_critical_at", [P, C.c_int], None)]:
    getattr(FJ, name).argtypes = args
    getattr(FJ, name).restype = res

L = G.lib()
L.cumf_jni_flatten_factors.argtypes = [P, P, C.c_int, C.c_int, P]
L.cumf_jni_flatten_factors.restype = C.c_int
L.cumf_jni_rows_to_array.argtypes = [P, C.c_int, C.c_int, P]
L.cumf_jni_rows_to_array.restype = P
JNI_DO = L.Java_org_apache_spark_ml_recommendation_CuMFJNIInterface_doALSWithCSR
JNI_DO.argtypes = [P, P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, P, P, P, P]
JNI_DO.restype = P


@pytest.fixture
def env():
    h = FJ.fj_create()
    yield h
    FJ.fj_destroy(h)


def exception(h):
    cls = C.create_string_buffer(256)
    msg = C.create_string_buffer(1024)
    if FJ.fj_exception(h, cls, 256, msg, 1024):
        return cls.value.decode(), msg.value.decode()
    return None


def factor_blocks(h, theta, sizes):
    """float[blocks][factors][f] Java object from a row-major theta."""
    blocks = FJ.fj_new_objects(h, len(sizes))
    k = 0
    for b, sz in enumerate(sizes):
        blk = FJ.fj_new_objects(h, sz)
        for j in range(sz):
            row = np.ascontiguousarray(theta[k], np.float32)
            FJ.fj_set(blk, j, FJ.fj_new_floats(h, row.ctypes.data_as(P), row.shape[0]))
            k += 1
        FJ.fj_set(blocks, b, blk)
    return blocks


def test_flatten_concatenates_blocks_in_order(env):
    theta = np.arange(7 * 3, dtype=np.float32).reshape(7, 3)
    blocks = factor_blocks(env, theta, [2, 0, 4, 1])
    out = np.full((7, 3), -1, np.float32)
    rc = L.cumf_jni_flatten_factors(FJ.fj_env(env), blocks, 3, 7, out.ctypes.data_as(P))
    assert rc == 0 and exception(env) is None
    np.testing.assert_array_equal(out, theta)
    assert FJ.fj_critical_depth(env) == 0


def test_flatten_rejects_wrong_count_and_length(env):
    theta = np.ones((4, 3), np.float32)
    blocks = factor_blocks(env, theta, [4])
    out = np.zeros((5, 3), np.float32)
    assert L.cumf_jni_flatten_factors(FJ.fj_env(env), blocks, 3, 5, out.ctypes.data_as(P)) == -1
    assert exception(env)[0] == "java/lang/IllegalArgumentException"
    FJ.fj_clear_exception(env)
    assert L.cumf_jni_flatten_factors(FJ.fj_env(env), blocks, 2, 4, out.ctypes.data_as(P)) == -1
    assert "wrong length" in exception(env)[1]
    assert FJ.fj_critical_depth(env) == 0


def test_flatten_reports_failed_critical_section(env):
    theta = np.ones((3, 2), np.float32)
    blocks = factor_blocks(env, theta, [3])
    FJ.fj_fail_critical_at(env, 1)
    out = np.zeros((3, 2), np.float32)
    assert L.cumf_jni_flatten_factors(FJ.fj_env(env), blocks, 2, 3, out.ctypes.data_as(P)) == -1
    assert exception(env)[0] == "java/lang/OutOfMemoryError"
    assert FJ.fj_critical_depth(env) == 0


def test_rows_to_array_layout(env):
    X = np.random.default_rng(0).standard_normal((5, 4)).astype(np.float32)
    arr = L.cumf_jni_rows_to_array(FJ.fj_env(env), 5, 4, X.ctypes.data_as(P))
    assert arr and FJ.fj_len(arr) == 5
    for i in range(5):
        row = np.zeros(4, np.float32)
        assert FJ.fj_floats(FJ.fj_get(arr, i), row.ctypes.data_as(P), 4) == 4
        np.testing.assert_array_equal(row, X[i])


def test_bridge_rejects_bad_lengths_before_touching_the_gpu(env):
    theta = np.ones((3, 2), np.float32)
    blocks = factor_blocks(env, theta, [3])
    rp = np.array([0, 1, 2], np.int32)  # m=2 -> needs 3 entries: OK
    ci = np.array([0, 1], np.int32)
    cv = np.array([1, 2], np.float32)
    jrp = FJ.fj_new_ints(env, rp.ctypes.data_as(P), 3)
    jci = FJ.fj_new_ints(env, ci.ctypes.data_as(P), 2)
    jcv = FJ.fj_new_floats(env, cv.ctypes.data_as(P), 2)
    # nnz=3 does not match the arrays
    res = JNI_DO(FJ.fj_env(env), None, 2, 3, 2, 3, 0.1, blocks, jrp, jci, jcv)
    assert res is None
    assert exception(env)[0] == "java/lang/IllegalArgumentException"
    assert FJ.fj_critical_depth(env) == 0


@pytest.mark.gpu
@pytest.mark.parametrize("f", [10, 100])
def test_bridge_end_to_end_matches_oracle(env, f):
    """The Spark entry point's success path: factor blocks flattened, the CSR
    pinned and uploaded, doALSWithCSR, float[m][f] back -- against the FP64
    oracle's half-sweep; called twice (the second call reuses the cached
    task plan) with every critical section released."""
    import cumf
    from oracle import oracle as O
    from tests import _data
    ds = _data.tiny()
    lam = 0.05
    theta = cumf.init_factors(ds.n, f, 3)
    sizes = [ds.n // 3, 0, ds.n - ds.n // 3 - 7, 7]
    Xo, _ = O.half_sweep(ds.csr_row_ptr, ds.csr_col, ds.csr_val, theta, lam)
    for call in range(2):
        blocks = factor_blocks(env, theta, sizes)
        rp = np.ascontiguousarray(ds.csr_row_ptr, np.int32)
        ci = np.ascontiguousarray(ds.csr_col, np.int32)
        cv = np.ascontiguousarray(ds.csr_val, np.float32)
        jrp = FJ.fj_new_ints(env, rp.ctypes.data_as(P), rp.shape[0])
        jci = FJ.fj_new_ints(env, ci.ctypes.data_as(P), ci.shape[0])
        jcv = FJ.fj_new_floats(env, cv.ctypes.data_as(P), cv.shape[0])
        res = JNI_DO(FJ.fj_env(env), None, ds.m, ds.n, f, ds.nnz, lam, blocks, jrp, jci, jcv)
        assert exception(env) is None
        assert res and FJ.fj_len(res) == ds.m
        assert FJ.fj_critical_depth(env) == 0
        X = np.zeros((ds.m, f), np.float32)
        for u in range(ds.m):
            assert FJ.fj_floats(FJ.fj_get(res, u), X[u].ctypes.data_as(P), f) == f
        rel = np.linalg.norm(X - Xo, axis=1) / np.maximum(np.linalg.norm(Xo, axis=1), 1e-30)
        assert rel.max() <= 1e-4, (call, rel.max())
