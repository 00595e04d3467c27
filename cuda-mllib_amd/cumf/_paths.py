"""Locations of the in-tree built libraries."""
import os

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REPO_DIR = os.path.dirname(PKG_DIR)
DATA_LIB = os.environ.get("CUMF_DATA_LIB") or os.path.join(PKG_DIR, "libcumf_data.so")
CPU_LIB = os.environ.get("CUMF_CPU_LIB") or os.path.join(PKG_DIR, "libcumf_cpu.so")
ALS_LIB = os.environ.get("CUMF_ALS_LIB") or os.path.join(PKG_DIR, "libGPUALS.so")
INCLUDE_DIR = os.path.join(REPO_DIR, "include")
