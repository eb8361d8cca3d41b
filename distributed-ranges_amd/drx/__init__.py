"""drx -- Python access to the MI355X sparse gemv library (libdrx_spmv.so).

The product is the C ABI in include/drx_spmv.h and the C++ host headers in
distributed-ranges_amd/include/dr/; this package is the ctypes shim used by
the autograd Function, the tests and the benchmark.
"""
from ._lib import (BF16, F16, F32, F64, I32, I64, DrxError, LIB_PATH, check, lib,
                   tune_get, tune_set, tuned)
from .core import CsrMatrix, Plan, Transpose, from_host
from . import gen
from .batch import Batch
from .mmio import mmread, read_csr_host
from .dense import gemm
from .autograd import SpMVFunction, spmv

__all__ = ["BF16", "F16", "F32", "F64", "I32", "I64", "DrxError", "LIB_PATH",
           "check", "lib", "tune_get", "tune_set", "tuned", "CsrMatrix", "Plan", "Transpose", "from_host", "gen",
           "SpMVFunction", "spmv", "Batch", "mmread", "read_csr_host", "gemm"]
