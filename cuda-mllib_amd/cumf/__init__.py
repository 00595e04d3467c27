"""cumf -- Python bindings over the MI355X cuMF ALS libraries.

The product is C++/HIP behind a C ABI (``libGPUALS.so``, declared in
``include/als.h``) and a host-only data generator (``libcumf_data.so``,
``include/cumf_data.h``).  This package only loads them with ctypes, for the
tests, the benchmark and Python callers; it computes nothing itself.
"""
from .data import Dataset, generate, init_factors, preset  # noqa: F401
