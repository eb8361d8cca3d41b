"""Synthetic inputs generated on the device (drx_gen_*)."""
import ctypes

import torch

from ._lib import check, lib
from .core import DTYPE_CODE, ITYPE_CODE, CsrMatrix, _p, _s


def uniform(n, dtype=torch.float32, lo=-1.0, hi=1.0, seed=0, device="cuda"):
    out = torch.empty(n, dtype=dtype, device=device)
    check(lib().drx_gen_uniform(DTYPE_CODE[dtype], _p(out), n, lo, hi, seed, _s()))
    return out


def normal(n, dtype=torch.float32, seed=0, device="cuda"):
    out = torch.empty(n, dtype=dtype, device=device)
    check(lib().drx_gen_normal(DTYPE_CODE[dtype], _p(out), n, seed, _s()))
    return out


def zipf_csr(nrows, ncols, s=1.5, lmax=4096, seed=0, vdtype=torch.float32,
             itype=torch.int32, block_nnz=0, device="cuda"):
    """Rows ~ Zipf(s) clipped to [1, lmax]; stratified sorted distinct
    columns; vals N(0,1)."""
    rowptr = torch.empty(nrows + 1, dtype=itype, device=device)
    nnz = ctypes.c_int64()
    check(lib().drx_gen_zipf_rowptr(ITYPE_CODE[itype], _p(rowptr), nrows, s,
                                    lmax, seed, ctypes.byref(nnz), _s()))
    nnz = nnz.value
    colind = torch.empty(nnz, dtype=itype, device=device)
    check(lib().drx_gen_stratified_colind(ITYPE_CODE[itype], _p(rowptr), nrows,
                                          ncols, _p(colind), seed + 1, _s()))
    vals = normal(nnz, vdtype, seed=seed + 2, device=device)
    return CsrMatrix(rowptr, colind, vals, (nrows, ncols), block_nnz=block_nnz)


def laplacian(dim, g, vdtype=torch.float64, itype=torch.int32, block_nnz=0,
              device="cuda"):
    nrows = g ** dim
    nnz = lib().drx_gen_laplacian_nnz(dim, g)
    rowptr = torch.empty(nrows + 1, dtype=itype, device=device)
    colind = torch.empty(nnz, dtype=itype, device=device)
    vals = torch.empty(nnz, dtype=vdtype, device=device)
    check(lib().drx_gen_laplacian(dim, g, ITYPE_CODE[itype], _p(rowptr),
                                  _p(colind), DTYPE_CODE[vdtype], _p(vals), _s()))
    return CsrMatrix(rowptr, colind, vals, (nrows, nrows), block_nnz=block_nnz)


def band(n, lo, hi, vdtype=torch.float32, itype=torch.int32, seed=0,
         block_nnz=0, device="cuda"):
    nnz = lib().drx_gen_band_nnz(n, lo, hi)
    rowptr = torch.empty(n + 1, dtype=itype, device=device)
    colind = torch.empty(nnz, dtype=itype, device=device)
    vals = torch.empty(nnz, dtype=vdtype, device=device)
    check(lib().drx_gen_band(n, lo, hi, ITYPE_CODE[itype], _p(rowptr),
                             _p(colind), DTYPE_CODE[vdtype], _p(vals), seed, _s()))
    return CsrMatrix(rowptr, colind, vals, (n, n), block_nnz=block_nnz)


def convert(src, dtype):
    out = torch.empty(src.shape, dtype=dtype, device=src.device)
    check(lib().drx_convert(DTYPE_CODE[src.dtype], _p(src), DTYPE_CODE[dtype],
                            _p(out), src.numel(), _s()))
    return out


def uniform_rows_csr(nrows, ncols, lmin, lmax, seed=0, vdtype=torch.float32,
                     itype=torch.int32, block_nnz=0, device="cuda"):
    """Row lengths uniform in [lmin, lmax]; stratified sorted distinct
    columns; vals N(0,1)."""
    rowptr = torch.empty(nrows + 1, dtype=itype, device=device)
    nnz = ctypes.c_int64()
    check(lib().drx_gen_uniform_rowptr(ITYPE_CODE[itype], _p(rowptr), nrows, lmin, lmax,
                                       seed, ctypes.byref(nnz), _s()))
    nnz = nnz.value
    colind = torch.empty(nnz, dtype=itype, device=device)
    check(lib().drx_gen_stratified_colind(ITYPE_CODE[itype], _p(rowptr), nrows, ncols,
                                          _p(colind), seed + 1, _s()))
    vals = normal(nnz, vdtype, seed=seed + 2, device=device)
    return CsrMatrix(rowptr, colind, vals, (nrows, ncols), block_nnz=block_nnz)


def gather(idx, x, out=None, stream=None):
    """out[i] = x[idx[i]] (drx_gather): int32 indices, 2/4/8-byte x."""
    if idx.dtype != torch.int32:
        raise ValueError("gather takes int32 indices")
    if out is None:
        out = torch.empty(idx.numel(), dtype=x.dtype, device=x.device)
    check(lib().drx_gather(_p(idx), idx.numel(), DTYPE_CODE[x.dtype], _p(x), x.numel(),
                           _p(out), _s(stream)))
    return out


def stream_read(buf, digest=None, stream=None):
    """Read every byte of the contiguous device tensor `buf` (drx_stream_read;
    the benchmark's HBM read-rate probe). Returns the digest tensor."""
    nbytes = buf.numel() * buf.element_size()
    if digest is None:
        digest = torch.empty(1 << 20, dtype=torch.int32, device=buf.device)
    check(lib().drx_stream_read(_p(buf), nbytes - nbytes % 16, _p(digest), digest.numel(),
                                _s(stream)))
    return digest
