"""Dense tile product on the matrix cores (drx_gemm).

gemm(a, b, c) is dr::sp::__detail::local_gemm
(include/dr/sp/algorithms/matrix/local_gemm.hpp:45-60): c += a @ b for
row-major views, here 2-D tensors whose rows are contiguous (stride(1) == 1)
and whose row stride is the leading dimension.
"""
import torch

from ._lib import check, lib
from .core import DTYPE_CODE, _p, _s


def _ld(t, name):
    if t.dim() != 2:
        raise ValueError(f"{name} must be 2-D")
    if t.size(1) > 1 and t.stride(1) != 1:
        raise ValueError(f"{name} rows must be contiguous")
    return max(t.stride(0), t.size(1), 1)


def gemm(a, b, c=None, alpha=1.0, beta=1.0, stream=None):
    """c = alpha * a @ b + beta * c (c is a new zeroed product when omitted);
    float32 or float64, all three of one dtype on one device."""
    if c is None:
        c = torch.empty(a.size(0), b.size(1), dtype=a.dtype, device=a.device)
        beta = 0.0
    if not (a.dtype == b.dtype == c.dtype) or a.dtype not in (torch.float32, torch.float64):
        raise ValueError("gemm takes float32 or float64 operands of one dtype")
    if a.size(1) != b.size(0) or c.shape != (a.size(0), b.size(1)):
        raise ValueError(f"shapes {tuple(a.shape)} @ {tuple(b.shape)} -> {tuple(c.shape)}")
    check(lib().drx_gemm(DTYPE_CODE[a.dtype], a.size(0), b.size(1), a.size(1), float(alpha),
                         _p(a), _ld(a, "a"), _p(b), _ld(b, "b"), float(beta), _p(c),
                         _ld(c, "c"), _s(stream)))
    return c
