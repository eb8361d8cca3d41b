"""torch.autograd.Function over the C ABI (gradient work only).

y = A(vals) x. Backward: dx = A^T dy (drx_spmv_adjoint over the transposed
structure, vals read through its permutation) and dvals[e] = dy[row(e)] *
x[col(e)] (drx_csr_dvals). Both run in libdrx_spmv.so.
"""
import torch

from .core import CsrMatrix


class SpMVFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vals, x, A: CsrMatrix):
        y = A.matvec(x, vals=vals)
        ctx.save_for_backward(vals, x)
        ctx.A = A
        return y

    @staticmethod
    def backward(ctx, dy):
        vals, x = ctx.saved_tensors
        A = ctx.A
        dy = dy.contiguous()
        dvals = dx = None
        if ctx.needs_input_grad[0]:
            xx = x if x.dtype == dy.dtype else x.to(dy.dtype)
            dvals = A.dvals(xx, dy).to(vals.dtype)
        if ctx.needs_input_grad[1]:
            dx = A.rmatvec(dy.to(torch.float64) if vals.dtype == torch.float64
                           else dy.to(torch.float32), vals=vals).to(x.dtype)
        return dvals, dx, None


def spmv(A: CsrMatrix, x, vals=None):
    """Differentiable y = A x with respect to x and A's values."""
    return SpMVFunction.apply(A.vals if vals is None else vals, x, A)
