"""drx -- Python access to the MI355X sparse gemv library (libdrx_spmv.so).

The product is the C ABI in include/drx_spmv.h and the C++ host headers in
distributed-ranges_amd/include/dr/; this package is the ctypes shim used by
the autograd Function, the tests and the benchmark.
"""
from ._lib import BF16, F16, F32, F64, I32, I64, DrxError, LIB_PATH, check, lib
from .core import CsrMatrix, Plan, Transpose, from_host
from . import gen
from .batch import Batch
from .autograd import SpMVFunction, spmv

__all__ = ["BF16", "F16", "F32", "F64", "I32", "I64", "DrxError", "LIB_PATH",
           "check", "lib", "CsrMatrix", "Plan", "Transpose", "from_host", "gen",
           "SpMVFunction", "spmv", "Batch"]
