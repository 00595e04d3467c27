"""The host task planner of the half-sweep (cumf_debug_plan; CPU only): every
rating is covered by exactly one task, tasks run longest first, split rows
own consecutive partial slots in near-equal chunks, and f > 140 plans whole
rows."""
import ctypes as C

import numpy as np
import pytest

from cumf import als as G
from tests import _data

L = G.lib()
L.cumf_debug_plan.argtypes = [C.c_void_p, C.c_int64, C.c_int, C.c_int, C.c_void_p, C.c_int64,
                              C.POINTER(C.c_int64), C.c_void_p, C.c_int64, C.POINTER(C.c_int64)]
L.cumf_debug_plan.restype = C.c_int


def plan(rp, split, f):
    rp = np.ascontiguousarray(rp, np.int32)
    m = rp.shape[0] - 1
    nt, nh = C.c_int64(), C.c_int64()
    assert L.cumf_debug_plan(rp.ctypes.data_as(C.c_void_p), m, split, f, None, 0, C.byref(nt), None, 0,
                             C.byref(nh)) in (0, 1)
    T = np.zeros((max(nt.value, 1), 4), np.int32)
    H = np.zeros((max(nh.value, 1), 4), np.int32)
    rc = L.cumf_debug_plan(rp.ctypes.data_as(C.c_void_p), m, split, f, T.ctypes.data_as(C.c_void_p), T.shape[0],
                           C.byref(nt), H.ctypes.data_as(C.c_void_p), H.shape[0], C.byref(nh))
    assert rc == 0, G.last_error()
    return T[:nt.value], H[:nh.value]


@pytest.mark.parametrize("split,f", [(256, 100), (64, 32), (16384, 100), (256, 141)])
def test_plan_covers_every_rating_once(split, f):
    ds = _data.netflix_like_small()
    rp = ds.csr_row_ptr
    T, H = plan(rp, split, f)
    deg = np.diff(rp)
    cover = np.zeros(ds.nnz, np.int32)
    for row, start, count, slot in T:
        assert rp[row] <= start and start + count <= rp[row + 1]
        cover[start:start + count] += 1
    assert np.all(cover == 1)
    # longest first, stable
    assert np.all(np.diff(T[:, 2]) <= 0)
    whole = T[T[:, 3] < 0]
    np.testing.assert_array_equal(np.sort(whole[:, 0]), np.flatnonzero(deg <= (split if f <= 140 else 1 << 62)))
    if f > 140:
        assert H.shape[0] == 0 and np.all(T[:, 3] == -1)
        return
    # split rows: consecutive slots, ceil(n / split) chunks of at most
    # `split` ratings whose sizes differ by at most one
    np.testing.assert_array_equal(H[:, 0], np.flatnonzero(deg > split))
    for row, slot0, nch, _ in H:
        mine = T[T[:, 0] == row]
        assert mine.shape[0] == nch == -(-deg[row] // split) and np.all(mine[:, 2] <= split)
        assert mine[:, 2].max() - mine[:, 2].min() <= 1
        np.testing.assert_array_equal(np.sort(mine[:, 3]), np.arange(slot0, slot0 + nch))
        assert mine[:, 2].sum() == deg[row]
    if H.shape[0]:
        assert H[0, 1] == 0 and np.all(H[1:, 1] == H[:-1, 1] + H[:-1, 2])


def test_plan_rejects_decreasing_row_pointers_and_small_buffers():
    rp = np.array([0, 3, 2, 5], np.int32)
    nt, nh = C.c_int64(), C.c_int64()
    assert L.cumf_debug_plan(rp.ctypes.data_as(C.c_void_p), 3, 16, 10, None, 0, C.byref(nt), None, 0,
                             C.byref(nh)) == 1
    assert "decrease" in G.last_error()
    rp = np.array([0, 3, 5], np.int32)
    assert L.cumf_debug_plan(rp.ctypes.data_as(C.c_void_p), 2, 16, 10, None, 0, C.byref(nt), None, 0,
                             C.byref(nh)) == 1
    assert nt.value == 2 and "too small" in G.last_error()
