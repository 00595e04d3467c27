"""The allocation-free half-sweep entry point (cumf_half_sweep_device_noalloc)
against the allocating one (cumf_half_sweep_device) through the C ABI: two
calls in a row on one GPU, the second on a smaller input, each compared with
the allocating entry point on the same device arrays."""
import ctypes as C

import numpy as np
import pytest

import cumf
from cumf import als as G
from tests import _data

pytestmark = pytest.mark.gpu

L = G.lib()
L.cumf_half_sweep_device_noalloc.argtypes = [C.c_void_p] * 6 + [C.c_int64, C.c_int64, C.c_int, C.c_int64, C.c_float]
L.cumf_half_sweep_device_noalloc.restype = C.c_int


def _sweep(noalloc, rp, ci, cv, theta, f, lam):
    m, n, nnz = rp.shape[0] - 1, theta.shape[0], ci.shape[0]
    bufs = [G.DeviceBuffer.from_array(a) for a in (rp, ci, cv, theta)]
    out = G.DeviceBuffer.from_array(np.full((m, f), np.nan, np.float32))
    try:
        if noalloc:
            rc = L.cumf_half_sweep_device_noalloc(None, *[b.ptr for b in bufs], out.ptr, m, n, f, nnz, lam)
            G._check(rc, "cumf_half_sweep_device_noalloc")  # enqueued on the null stream; the
            # synchronous copy below is ordered after it
        else:
            opt = G.default_options()
            G._check(L.cumf_half_sweep_device(None, *[b.ptr for b in bufs], out.ptr, m, n, f, nnz, lam, C.byref(opt)),
                     "cumf_half_sweep_device")
        return out.to_array(np.float32, m * f).reshape(m, f)
    finally:
        for b in bufs + [out]:
            b.free()


def _case(ds, rows, f, lam, seed):
    rp = np.ascontiguousarray(ds.csr_row_ptr[:rows + 1], np.int32)
    nnz = int(rp[-1])
    ci = np.ascontiguousarray(ds.csr_col[:nnz], np.int32)
    cv = np.ascontiguousarray(ds.csr_val[:nnz], np.float32)
    theta = cumf.init_factors(ds.n, f, seed)
    return rp, ci, cv, theta


@pytest.mark.parametrize("f", [32, 100])
def test_noalloc_matches_allocating_entry_point_twice(f):
    ds = _data.netflix_like_small()
    lam = 0.058
    for rows, seed in ((ds.m, 21), (ds.m // 3, 22)):  # the second call on a smaller input
        rp, ci, cv, theta = _case(ds, rows, f, lam, seed)
        Xa = _sweep(False, rp, ci, cv, theta, f, lam)
        Xn = _sweep(True, rp, ci, cv, theta, f, lam)
        assert np.all(np.isfinite(Xn))
        # rows below the allocating path's split size (16384 ratings; every
        # row here) run the same arithmetic: bit-identical
        np.testing.assert_array_equal(Xn, Xa)


def test_noalloc_rejects_unsupported_shapes():
    rp = np.array([0, 1], np.int32)
    for f in (30, 144, 0):
        rc = L.cumf_half_sweep_device_noalloc(None, rp.ctypes.data, rp.ctypes.data, rp.ctypes.data, rp.ctypes.data,
                                              rp.ctypes.data, 1, 1, f, 1, 0.1)
        assert rc == 1 and "multiple of 4" in G.last_error()
