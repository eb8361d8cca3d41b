"""Matrix Market input: the library's parallel host reader (drx_mm_read_csr)
and upload to a device CsrMatrix.

Mirrors dr::read_csr<T, I>(path, one_indexed)
(include/dr/detail/matrix_io.hpp:243-252) and, with from_host, the input
step of dr::sp::mmread (include/dr/sp/util/matrix_io.hpp:68-84).
"""
import ctypes

import numpy as np
import torch

from ._lib import F32, F64, I32, I64, check, lib

_NP_I = {I32: np.int32, I64: np.int64}
_NP_V = {F64: np.float64, F32: np.float32}


def read_csr_host(path, dtype=np.float64, itype=np.int64, one_indexed=True, nthreads=0):
    """Host CSR arrays (rowptr, colind, vals) and the shape (m, n) of a
    coordinate Matrix Market file."""
    ic = I32 if np.dtype(itype) == np.int32 else I64
    vc = F32 if np.dtype(dtype) == np.float32 else F64
    m, n, nnz = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    rp, ci, v = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    L = lib()
    check(L.drx_mm_read_csr(str(path).encode(), 1 if one_indexed else 0, ic, vc,
                            int(nthreads), ctypes.byref(m), ctypes.byref(n),
                            ctypes.byref(nnz), ctypes.byref(rp), ctypes.byref(ci),
                            ctypes.byref(v)))
    try:
        def take(ptr, count, npdt):
            if count == 0:
                return np.zeros(0, npdt)
            buf = (ctypes.c_char * (count * np.dtype(npdt).itemsize)).from_address(ptr.value)
            return np.frombuffer(buf, dtype=npdt).copy()
        rowptr = take(rp, m.value + 1, _NP_I[ic])
        colind = take(ci, nnz.value, _NP_I[ic])
        vals = take(v, nnz.value, _NP_V[vc])
    finally:
        for ptr in (rp, ci, v):
            L.drx_host_free(ptr)
    return rowptr, colind, vals, (m.value, n.value)


def mmread(path, dtype=torch.float64, itype=torch.int32, one_indexed=True, nthreads=0,
           device="cuda", block_nnz=0):
    """Read a Matrix Market file into a device CsrMatrix."""
    from .core import from_host
    npv = np.float32 if dtype == torch.float32 else np.float64
    npi = np.int32 if itype == torch.int32 else np.int64
    rowptr, colind, vals, shape = read_csr_host(path, npv, npi, one_indexed, nthreads)
    return from_host(rowptr, colind, vals, shape, device=device, itype=itype,
                     block_nnz=block_nnz)
