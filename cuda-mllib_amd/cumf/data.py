"""ctypes binding of libcumf_data.so (include/cumf_data.h).

Produces seeded synthetic rating matrices as numpy arrays: training CSR and
CSC (int32 indices, float32 values) and a held-out COO set.
"""
import ctypes as C
import dataclasses

import numpy as np

from . import _paths


class GenParams(C.Structure):
    _fields_ = [
        ("m", C.c_int64), ("n", C.c_int64), ("nnz_total", C.c_int64), ("n_holdout", C.c_int64),
        ("row_s", C.c_double), ("col_s", C.c_double), ("row_cap", C.c_int64), ("col_cap", C.c_int64),
        ("rank", C.c_int32), ("latent_var", C.c_double), ("noise_std", C.c_double), ("mu", C.c_double),
        ("round_clip", C.c_int32), ("seed", C.c_uint64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class GenInfo(C.Structure):
    _fields_ = [
        ("m", C.c_int64), ("n", C.c_int64), ("nnz", C.c_int64), ("n_test", C.c_int64),
        ("q_row", C.c_double), ("q_col", C.c_double),
        ("max_row_deg", C.c_int64), ("min_row_deg", C.c_int64),
        ("max_col_deg", C.c_int64), ("min_col_deg", C.c_int64),
        ("hist", C.c_int64 * 5),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(_paths.DATA_LIB)
        P = C.c_void_p
        L.cumf_gen_preset.argtypes = [C.c_char_p, C.POINTER(GenParams)]
        L.cumf_gen_create.argtypes = [C.POINTER(GenParams), C.POINTER(P)]
        L.cumf_gen_info.argtypes = [P, C.POINTER(GenInfo)]
        L.cumf_gen_copy.argtypes = [P] + [P] * 9
        L.cumf_gen_free.argtypes = [P]
        L.cumf_gen_free.restype = None
        L.cumf_gen_init_factors.argtypes = [C.c_int64, C.c_int, C.c_uint64, P]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


@dataclasses.dataclass
class Dataset:
    m: int
    n: int
    csr_row_ptr: np.ndarray
    csr_col: np.ndarray
    csr_val: np.ndarray
    csc_col_ptr: np.ndarray
    csc_row: np.ndarray
    csc_val: np.ndarray
    test_u: np.ndarray
    test_v: np.ndarray
    test_r: np.ndarray
    params: dict
    info: dict

    @property
    def nnz(self):
        return int(self.csr_col.shape[0])


def preset(name, **overrides):
    p = GenParams()
    if lib().cumf_gen_preset(name.encode(), C.byref(p)) != 0:
        raise ValueError(f"unknown workload preset {name!r}")
    for k, v in overrides.items():
        setattr(p, k, v)
    return p


def generate(params):
    """Runs the generator; returns a Dataset of numpy arrays."""
    L = lib()
    h = C.c_void_p()
    if L.cumf_gen_create(C.byref(params), C.byref(h)) != 0:
        raise ValueError("cumf_gen_create rejected the parameters")
    try:
        info = GenInfo()
        L.cumf_gen_info(h, C.byref(info))
        m, n, nnz, nt = info.m, info.n, info.nnz, info.n_test
        a = dict(
            csr_row_ptr=np.empty(m + 1, np.int32), csr_col=np.empty(nnz, np.int32),
            csr_val=np.empty(nnz, np.float32), csc_col_ptr=np.empty(n + 1, np.int32),
            csc_row=np.empty(nnz, np.int32), csc_val=np.empty(nnz, np.float32),
            test_u=np.empty(nt, np.int32), test_v=np.empty(nt, np.int32), test_r=np.empty(nt, np.float32))
        L.cumf_gen_copy(h, *[_ptr(a[k]) for k in (
            "csr_row_ptr", "csr_col", "csr_val", "csc_col_ptr", "csc_row", "csc_val",
            "test_u", "test_v", "test_r")])
        inf = {k: getattr(info, k) for k, _ in GenInfo._fields_ if k != "hist"}
        inf["hist"] = list(info.hist)
        return Dataset(m=m, n=n, params=params.as_dict(), info=inf, **a)
    finally:
        L.cumf_gen_free(h)


def init_factors(rows, f, seed):
    """rows x f float32 factors, each row N(0,1)^f scaled to unit norm."""
    out = np.empty((rows, f), np.float32)
    if lib().cumf_gen_init_factors(rows, f, seed, _ptr(out)) != 0:
        raise ValueError("bad factor shape")
    return out
