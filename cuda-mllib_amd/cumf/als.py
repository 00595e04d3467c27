"""ctypes binding of libGPUALS.so (include/als.h, include/utilities.h).

``ALS`` wraps the cumf_* context API (ratings and factors resident on the
device, whole iterations, held-out RMSE, per-kernel timing).
``do_als_with_csr`` drives the drop-in entry point doALSWithCSR exactly as
the reference's JNI bridge does (als/src/CuMFJNIInterface.cpp:89-109):
CSR arrays in mallocBest memory, consumed by the call.

The library is required: every function raises if libGPUALS.so is missing
or reports an error; there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

from . import _paths

CUMF_OPTIONS_VERSION = 1
CUMF_PREC_FP32 = 0
CUMF_PREC_BF16_GRAM = 1
CUMF_SOLVER_CHOLESKY = 0
CUMF_SOLVER_CG = 1
CUMF_WARM_PREV = 0
CUMF_WARM_ZERO = 1
CUMF_WARM_MLP = 2
CUMF_WARM_BEST = 3


class Options(C.Structure):
    _fields_ = [("version", C.c_int32), ("precision", C.c_int32), ("split", C.c_int32),
                ("timing", C.c_int32), ("use_graph", C.c_int32),
                ("compat_hermitian_bug", C.c_int32), ("solver", C.c_int32),
                ("cg_iters", C.c_int32), ("cg_rounds", C.c_int32), ("cg_tol", C.c_float),
                ("warm_start", C.c_int32), ("warm_sweeps", C.c_int32), ("reserved", C.c_int32 * 4)]


class Stats(C.Structure):
    _fields_ = [("task_ms", C.c_double), ("task_launches", C.c_int64),
                ("finish_ms", C.c_double), ("finish_launches", C.c_int64),
                ("rmse_ms", C.c_double), ("rmse_launches", C.c_int64),
                ("sweep_ms", C.c_double), ("sweeps", C.c_int64),
                ("failed_rows", C.c_int64), ("gram_flops", C.c_int64),
                ("iter_ms", C.c_double), ("iters", C.c_int64),
                ("cg_fallback_rows", C.c_int64), ("warm_ms", C.c_double), ("warm_launches", C.c_int64),
                ("fallback_ms", C.c_double), ("fallback_launches", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class CumfError(RuntimeError):
    pass


_lib = None
P = C.c_void_p
I32P = C.POINTER(C.c_int32)


def lib():
    """Loads libGPUALS.so (raises OSError if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_paths.ALS_LIB):
        raise OSError(f"{_paths.ALS_LIB} is not built; run __graft_entry__.build()")
    L = C.CDLL(_paths.ALS_LIB)
    sig = {
        "cumf_default_options": ([C.POINTER(Options)], C.c_int),
        "cumf_create": ([C.c_int, C.c_int, C.c_float, C.POINTER(Options), C.POINTER(P)], C.c_int),
        "cumf_destroy": ([P], C.c_int),
        "cumf_set_ratings": ([P, C.c_int64, C.c_int64, C.c_int64, P, P, P, P, P, P], C.c_int),
        "cumf_set_heldout": ([P, C.c_int64, P, P, P], C.c_int),
        "cumf_set_factors": ([P, P, P], C.c_int),
        "cumf_get_factors": ([P, P, P], C.c_int),
        "cumf_update_x": ([P], C.c_int),
        "cumf_update_theta": ([P], C.c_int),
        "cumf_iterate": ([P, C.c_int], C.c_int),
        "cumf_heldout_rmse": ([P, C.POINTER(C.c_double)], C.c_int),
        "cumf_synchronize": ([P], C.c_int),
        "cumf_get_stats": ([P, C.POINTER(Stats)], C.c_int),
        "cumf_reset_stats": ([P], C.c_int),
        "cumf_get_stream": ([P], P),
        "cumf_last_error": ([], C.c_char_p),
        "cumf_half_sweep_device": ([P, P, P, P, P, P, C.c_int64, C.c_int64, C.c_int, C.c_int64,
                                    C.c_float, C.POINTER(Options)], C.c_int),
        "cumf_debug_gram_dim": ([C.c_int], C.c_int),
        "cumf_debug_row_gram": ([P, C.c_int, P, C.c_int64, P], C.c_int),
        "cumf_debug_mfma_layout": ([C.c_int, P], C.c_int),
        "cumf_debug_mfma_bf16_layout": ([C.c_int, P], C.c_int),
        "cumf_debug_compat_system": ([P, C.c_int, P, C.c_int64, P], C.c_int),
        "cumf_debug_copy": ([C.c_int, C.c_int64, C.c_int, C.POINTER(C.c_double)], C.c_int),
        "cumf_debug_force_host_alloc": ([C.c_int], C.c_int),
        "cumf_debug_large_dual": ([C.c_int], C.c_int),
        "cumf_debug_cg_trace": ([P, C.c_int, P], C.c_int),
        "cumf_debug_features": ([P, C.c_int, P, C.c_int64, P], C.c_int),
        "cumf_set_mlp": ([P, P], C.c_int),
        "cumf_load_tile_table": ([P, C.c_char_p], C.c_int),
        "cumf_save_tile_table": ([P, C.c_char_p], C.c_int),
        "cumf_autotune": ([P, C.c_int, C.c_char_p], C.c_int),
        "cumf_mlp_create": ([C.c_int, C.c_int, C.c_uint64, C.POINTER(P)], C.c_int),
        "cumf_mlp_destroy": ([P], C.c_int),
        "cumf_mlp_dims": ([P, P, C.POINTER(C.c_int64)], C.c_int),
        "cumf_mlp_get_params": ([P, P], C.c_int),
        "cumf_mlp_set_params": ([P, P], C.c_int),
        "cumf_mlp_predict_host": ([P, P, C.c_int64, P], C.c_int),
        "cumf_mlp_gradients_host": ([P, P, P, C.c_int64, P, C.POINTER(C.c_double)], C.c_int),
        "cumf_mlp_step_host": ([P, P, P, C.c_int64, C.c_float, C.POINTER(C.c_double)], C.c_int),
        "cumf_mlp_harvest": ([P, P, C.c_int], C.c_int),
        "cumf_mlp_harvested": ([P], C.c_int64),
        "cumf_mlp_clear_harvest": ([P], C.c_int),
        "cumf_mlp_train": ([P, C.c_int, C.c_int, C.c_float, C.c_uint64, C.POINTER(C.c_double)], C.c_int),
        "cumf_memcpy_h2d": ([P, P, C.c_size_t], C.c_int),
        "cumf_memcpy_d2h": ([P, P, C.c_size_t], C.c_int),
        "cumf_malloc_best": ([C.POINTER(P), C.c_size_t], C.c_int),
        "cumf_free_best": ([P], C.c_int),
        "get_gpu": ([], C.c_int),
        "cumf_load_csr_bin": ([C.c_char_p, C.c_char_p, C.c_char_p, P, P, P, C.c_int, C.c_long], C.c_int),
        "cumf_save_csr_bin": ([C.c_char_p, C.c_char_p, C.c_char_p, P, P, P, C.c_int, C.c_long], C.c_int),
        "cumf_split_holdout": ([C.c_int64, C.c_int64, C.c_int64, P, P, P, C.c_double, C.c_uint64, P, P, P,
                                C.POINTER(C.c_int64), P, P, P, C.POINTER(C.c_int64)], C.c_int),
        "doALSWithCSR": ([P, P, P, P, P, P, C.c_int, C.c_int, C.c_int, C.c_long, C.c_float, C.c_int],
                         None),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name, None)
        if fn is None:  # entry point absent from the loaded library: calling it raises
            continue
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def last_error():
    return lib().cumf_last_error().decode()


def _check(rc, what):
    if rc != 0:
        raise CumfError(f"{what} failed (status {rc}): {last_error()}")


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def default_options(**kw):
    o = Options()
    _check(lib().cumf_default_options(C.byref(o)), "cumf_default_options")
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class ALS:
    """Resident ALS state on one GPU (cumf_* context API)."""

    def __init__(self, f, lam, device=0, **opts):
        self.f = int(f)
        self.lam = float(lam)
        self._h = P()
        self.opt = default_options(**opts)
        _check(lib().cumf_create(device, self.f, self.lam, C.byref(self.opt), C.byref(self._h)),
               "cumf_create")
        self.m = self.n = None

    def close(self):
        if self._h:
            lib().cumf_destroy(self._h)
            self._h = P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_ratings(self, m, n, row_ptr, col, val, csc_col_ptr=None, csc_row=None, csc_val=None):
        arrs = [np.ascontiguousarray(row_ptr, np.int32), np.ascontiguousarray(col, np.int32),
                np.ascontiguousarray(val, np.float32)]
        csc = [None if a is None else np.ascontiguousarray(a, t)
               for a, t in ((csc_col_ptr, np.int32), (csc_row, np.int32), (csc_val, np.float32))]
        nnz = arrs[1].shape[0]
        _check(lib().cumf_set_ratings(self._h, m, n, nnz, *[_ptr(a) for a in arrs],
                                      *[_ptr(a) for a in csc]), "cumf_set_ratings")
        self.m, self.n = int(m), int(n)

    def set_dataset(self, ds):
        self.set_ratings(ds.m, ds.n, ds.csr_row_ptr, ds.csr_col, ds.csr_val,
                         ds.csc_col_ptr, ds.csc_row, ds.csc_val)
        self.set_heldout(ds.test_u, ds.test_v, ds.test_r)

    def set_heldout(self, u, v, r):
        u = np.ascontiguousarray(u, np.int32)
        v = np.ascontiguousarray(v, np.int32)
        r = np.ascontiguousarray(r, np.float32)
        _check(lib().cumf_set_heldout(self._h, u.shape[0], _ptr(u), _ptr(v), _ptr(r)),
               "cumf_set_heldout")

    def set_factors(self, X=None, theta=None):
        X = None if X is None else np.ascontiguousarray(X, np.float32)
        theta = None if theta is None else np.ascontiguousarray(theta, np.float32)
        if X is not None:
            assert X.shape == (self.m, self.f)
        if theta is not None:
            assert theta.shape == (self.n, self.f)
        _check(lib().cumf_set_factors(self._h, _ptr(X), _ptr(theta)), "cumf_set_factors")

    def get_factors(self):
        X = np.empty((self.m, self.f), np.float32)
        T = np.empty((self.n, self.f), np.float32)
        _check(lib().cumf_get_factors(self._h, _ptr(X), _ptr(T)), "cumf_get_factors")
        return X, T

    def update_x(self):
        _check(lib().cumf_update_x(self._h), "cumf_update_x")

    def update_theta(self):
        _check(lib().cumf_update_theta(self._h), "cumf_update_theta")

    def iterate(self, iters=1):
        _check(lib().cumf_iterate(self._h, iters), "cumf_iterate")

    def rmse(self):
        out = C.c_double()
        _check(lib().cumf_heldout_rmse(self._h, C.byref(out)), "cumf_heldout_rmse")
        return out.value

    def synchronize(self):
        _check(lib().cumf_synchronize(self._h), "cumf_synchronize")

    def stats(self):
        s = Stats()
        _check(lib().cumf_get_stats(self._h, C.byref(s)), "cumf_get_stats")
        return s.as_dict()

    def reset_stats(self):
        _check(lib().cumf_reset_stats(self._h), "cumf_reset_stats")

    def compat_system(self, side, rows):
        """[M | b] the compat solver sees for listed rows (f x (f+1), FP64)."""
        rows = np.ascontiguousarray(rows, np.int32)
        out = np.empty((rows.shape[0], self.f * (self.f + 1)), np.float64)
        _check(lib().cumf_debug_compat_system(self._h, side, _ptr(rows), rows.shape[0], _ptr(out)),
               "cumf_debug_compat_system")
        M = out[:, :self.f * self.f].reshape(-1, self.f, self.f)
        return M, out[:, self.f * self.f:]

    def cg_trace(self, side):
        """One half-sweep (side 0 = X, 1 = Theta) recording per row: initial and
        final relative residual, refinement rounds, fell back (0/1)."""
        rows = self.m if side == 0 else self.n
        out = np.zeros((rows, 4), np.float32)
        _check(lib().cumf_debug_cg_trace(self._h, side, _ptr(out)), "cumf_debug_cg_trace")
        return out

    def features(self, side, rows):
        """FP32 values of the BF16 warm-start features of listed rows."""
        rows = np.ascontiguousarray(rows, np.int32)
        k0 = (2 * self.f + 1 + 31) // 32 * 32
        out = np.empty((rows.shape[0], k0), np.float32)
        _check(lib().cumf_debug_features(self._h, side, _ptr(rows), rows.shape[0], _ptr(out)),
               "cumf_debug_features")
        return out

    def autotune(self, reps=3, path=None):
        """Times every degree bucket's launch shapes and keeps the fastest;
        writes the tile table to path if given."""
        _check(lib().cumf_autotune(self._h, int(reps), None if path is None else str(path).encode()),
               "cumf_autotune")

    def load_tile_table(self, path):
        _check(lib().cumf_load_tile_table(self._h, str(path).encode()), "cumf_load_tile_table")

    def save_tile_table(self, path):
        _check(lib().cumf_save_tile_table(self._h, str(path).encode()), "cumf_save_tile_table")

    def set_mlp(self, mlp):
        _check(lib().cumf_set_mlp(self._h, None if mlp is None else mlp._h), "cumf_set_mlp")
        self._mlp = mlp  # keeps the network alive while attached

    def row_gram(self, side, rows):
        """Augmented Gram (D x D, ridge included) of listed rows; side 0 = X."""
        rows = np.ascontiguousarray(rows, np.int32)
        D = lib().cumf_debug_gram_dim(self.f)
        out = np.empty((rows.shape[0], D, D), np.float32)
        _check(lib().cumf_debug_row_gram(self._h, side, _ptr(rows), rows.shape[0], _ptr(out)),
               "cumf_debug_row_gram")
        return out


class DeviceBuffer:
    """mallocBest allocation (device, or mapped host on fallback)."""

    def __init__(self, nbytes):
        self.ptr = P()
        _check(lib().cumf_malloc_best(C.byref(self.ptr), max(int(nbytes), 4)), "mallocBest")
        self.nbytes = int(nbytes)

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        if a.nbytes:
            _check(lib().cumf_memcpy_h2d(b.ptr, _ptr(a), a.nbytes), "cumf_memcpy_h2d")
        return b

    def to_array(self, dtype, count):
        out = np.empty(count, dtype)
        if out.nbytes:
            _check(lib().cumf_memcpy_d2h(_ptr(out), self.ptr, out.nbytes), "cumf_memcpy_d2h")
        return out

    def free(self):
        if self.ptr:
            _check(lib().cumf_free_best(self.ptr), "freeBest")
            self.ptr = P()


def do_als_with_csr(row_ptr, col, val, theta, m, n, f, lam, x_batch=1):
    """One half-sweep through doALSWithCSR; returns X (m x f).

    Uploads the CSR into mallocBest memory, which the call consumes (the
    buffers must not be freed again), like the reference's JNI bridge.
    """
    L = lib()
    theta = np.ascontiguousarray(theta, np.float32)
    if f > 0:
        theta = theta.reshape(n, f)
    X = np.zeros((m, max(int(f), 1)), np.float32)
    bufs = [DeviceBuffer.from_array(np.asarray(row_ptr, np.int32)),
            DeviceBuffer.from_array(np.asarray(col, np.int32)),
            DeviceBuffer.from_array(np.asarray(val, np.float32))]
    nnz = int(np.asarray(col).shape[0])
    L.doALSWithCSR(None, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, _ptr(theta), _ptr(X), m, n, f, nnz,
                   lam, x_batch)
    for b in bufs:  # consumed by the call
        b.ptr = P()
    err = last_error()
    if err:
        raise CumfError(f"doALSWithCSR failed: {err}")
    return X


def mfma_layout(device=0):
    out = np.empty((64, 4), np.float32)
    _check(lib().cumf_debug_mfma_layout(device, _ptr(out)), "cumf_debug_mfma_layout")
    return out


def mfma_bf16_layout(device=0):
    out = np.empty((64, 4), np.float32)
    _check(lib().cumf_debug_mfma_bf16_layout(device, _ptr(out)), "cumf_debug_mfma_bf16_layout")
    return out


def copy_kernel_ms(nbytes, iters=20, device=0):
    ms = C.c_double()
    _check(lib().cumf_debug_copy(device, int(nbytes), iters, C.byref(ms)), "cumf_debug_copy")
    return ms.value


def save_csr_bin(prefix, row_ptr, col, val):
    """Writes <prefix>.row.bin / .col.bin / .data.bin (raw int32/int32/fp32)."""
    rp = np.ascontiguousarray(row_ptr, np.int32)
    ci = np.ascontiguousarray(col, np.int32)
    cv = np.ascontiguousarray(val, np.float32)
    _check(lib().cumf_save_csr_bin(f"{prefix}.data.bin".encode(), f"{prefix}.row.bin".encode(),
                                   f"{prefix}.col.bin".encode(), _ptr(cv), _ptr(rp), _ptr(ci),
                                   rp.shape[0] - 1, ci.shape[0]), "cumf_save_csr_bin")


def load_csr_bin(prefix, m, nnz):
    """Reads the raw CSR files written by save_csr_bin (or the reference's tools)."""
    rp = np.empty(m + 1, np.int32)
    ci = np.empty(nnz, np.int32)
    cv = np.empty(nnz, np.float32)
    _check(lib().cumf_load_csr_bin(f"{prefix}.data.bin".encode(), f"{prefix}.row.bin".encode(),
                                   f"{prefix}.col.bin".encode(), _ptr(cv), _ptr(rp), _ptr(ci), m, nnz),
           "cumf_load_csr_bin")
    return rp, ci, cv


def split_holdout(row_ptr, col, val, n, fraction, seed=0):
    """Seeded held-out split (cumf_split_holdout): returns the training CSR
    (row_ptr, col, val) and the held-out COO (u, v, r)."""
    rp = np.ascontiguousarray(row_ptr, np.int32)
    ci = np.ascontiguousarray(col, np.int32)
    cv = np.ascontiguousarray(val, np.float32)
    m, nnz = rp.shape[0] - 1, ci.shape[0]
    trp = np.empty(m + 1, np.int32)
    tci, tcv = np.empty(nnz, np.int32), np.empty(nnz, np.float32)
    hu, hv, hr = np.empty(nnz, np.int32), np.empty(nnz, np.int32), np.empty(nnz, np.float32)
    nt, nh = C.c_int64(), C.c_int64()
    _check(lib().cumf_split_holdout(m, int(n), nnz, _ptr(rp), _ptr(ci), _ptr(cv), float(fraction), int(seed),
                                    _ptr(trp), _ptr(tci), _ptr(tcv), C.byref(nt), _ptr(hu), _ptr(hv), _ptr(hr),
                                    C.byref(nh)), "cumf_split_holdout")
    a, b = nt.value, nh.value
    return trp, tci[:a].copy(), tcv[:a].copy(), hu[:b].copy(), hv[:b].copy(), hr[:b].copy()
