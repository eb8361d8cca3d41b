"""Device-side CSR objects over the C ABI, with torch tensors as device
memory (torch is plumbing here: allocation, streams, events).

``CsrMatrix`` holds one CSR share in HBM (rowptr/colind/vals), its plan
(row blocks for the kernel) and, on demand, its transposed structure for the
adjoint. All math runs in libdrx_spmv.so; nothing here computes on the CPU.
"""
import ctypes

import torch

from ._lib import BF16, F16, F32, F64, I32, I64, DrxError, check, lib

DTYPE_CODE = {torch.float64: F64, torch.float32: F32, torch.bfloat16: BF16,
              torch.float16: F16}
CODE_DTYPE = {v: k for k, v in DTYPE_CODE.items()}
ITYPE_CODE = {torch.int32: I32, torch.int64: I64}


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def _s(stream=None):
    if stream is None:
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


def acc_dtype(vdtype):
    return torch.float64 if vdtype == torch.float64 else torch.float32


class Plan:
    """Row-block analysis of one CSR share (drx_plan_create)."""

    def __init__(self, rowptr, nrows, ncols, nnz, base=0, has_sentinel=True,
                 block_nnz=0, stream=None):
        if rowptr.device.type != "cuda":
            raise ValueError("rowptr must be a device tensor")
        self.rowptr = rowptr  # kept alive: the plan references it
        self.itype = ITYPE_CODE[rowptr.dtype]
        self.nrows, self.ncols, self.nnz = int(nrows), int(ncols), int(nnz)
        self.base = int(base)
        self.has_sentinel = bool(has_sentinel)
        h = ctypes.c_void_p()
        check(lib().drx_plan_create(ctypes.byref(h), self.itype, _p(rowptr),
                                    self.base, self.nrows, self.ncols,
                                    self.nnz, int(has_sentinel),
                                    int(block_nnz), _s(stream)))
        self.handle = h

    def info(self):
        nb, bn = ctypes.c_int64(), ctypes.c_int64()
        check(lib().drx_plan_info(self.handle, ctypes.byref(nb), ctypes.byref(bn)))
        return nb.value, bn.value

    def close(self):
        if getattr(self, "handle", None):
            lib().drx_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Transpose:
    """CSC of a share (drx_transpose_create) for the adjoint."""

    def __init__(self, plan, colind, stream=None):
        self.plan = plan
        h = ctypes.c_void_p()
        check(lib().drx_transpose_create(ctypes.byref(h), plan.handle,
                                         _p(colind), _s(stream)))
        self.handle = h

    def arrays(self):
        """(colptr, rowind, perm) copied to host numpy arrays (for tests)."""
        cp, ri, pm = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        check(lib().drx_transpose_arrays(self.handle, ctypes.byref(cp),
                                         ctypes.byref(ri), ctypes.byref(pm)))
        idt = torch.int32 if self.plan.itype == I32 else torch.int64
        out = []
        for ptr, n in ((cp, self.plan.ncols + 1), (ri, self.plan.nnz),
                       (pm, self.plan.nnz)):
            t = torch.empty(n, dtype=idt, device="cuda")
            if n:
                torch.cuda.synchronize()
                _copy_dev(t, ptr.value, n * t.element_size())
            out.append(t)
        return out

    def close(self):
        if getattr(self, "handle", None):
            lib().drx_transpose_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _copy_dev(dst, src_ptr, nbytes):
    """Copy from a raw device pointer into tensor dst (drx_copy_device)."""
    check(lib().drx_copy_device(_p(dst), ctypes.c_void_p(src_ptr), nbytes, _s()))


class CsrMatrix:
    """One CSR share resident in HBM.

    rowptr: nrows (+1 with sentinel) absolute offsets; colind/vals: the
    share's nnz entries; base: absolute offset of entry 0 of the share.
    """

    def __init__(self, rowptr, colind, vals, shape, base=0, has_sentinel=True,
                 block_nnz=0, stream=None):
        self.rowptr, self.colind, self.vals = rowptr, colind, vals
        self.shape = (int(shape[0]), int(shape[1]))
        self.nnz = int(colind.numel())
        self.device = rowptr.device
        if colind.dtype != rowptr.dtype:
            raise ValueError("rowptr and colind must share an index dtype")
        if colind.device != self.device or (colind.dim() != 1 or
                                             (self.nnz > 1 and colind.stride(0) != 1)):
            raise ValueError("colind must be a contiguous 1-D tensor on rowptr's device")
        if not rowptr.is_contiguous():
            raise ValueError("rowptr must be contiguous")
        _vals_ok(vals, self.nnz, self.device)
        self.base = int(base)
        self.has_sentinel = bool(has_sentinel)
        self.plan = Plan(rowptr, self.shape[0], self.shape[1], self.nnz,
                         base=base, has_sentinel=has_sentinel,
                         block_nnz=block_nnz, stream=stream)
        self._t = None
        self._tvals_key = None  # (tensor id, version) of the values stored in CSC order
        self._tvals_ref = None

    @property
    def vcode(self):
        return DTYPE_CODE[self.vals.dtype]

    def matvec(self, x, y=None, alpha=1.0, beta=0.0, vals=None, stream=None):
        """y = alpha*A*x + beta*y; x is (n,) or column-major (n, k) given as
        a (k, n) contiguous tensor (row j = right-hand side j)."""
        vals = self.vals if vals is None else _vals_ok(vals, self.nnz, self.device)
        k, ldx = _cols(_on(x, self.device, "x"), self.shape[1])
        ydt = torch.float64 if vals.dtype == torch.float64 else torch.float32
        if y is None:
            shape = (self.shape[0],) if x.dim() == 1 else (k, self.shape[0])
            y = torch.empty(shape, dtype=ydt, device=x.device)
        ky, ldy = _cols(_on(y, self.device, "y"), self.shape[0])
        if ky != k:
            raise ValueError("x and y differ in their number of right-hand sides")
        check(lib().drx_spmv(self.plan.handle, DTYPE_CODE[vals.dtype],
                             _p(self.colind), _p(vals), DTYPE_CODE[x.dtype],
                             _p(x), ldx, DTYPE_CODE[y.dtype], _p(y), ldy, k,
                             float(alpha), float(beta), _s(stream)))
        return y

    def set_panels(self, panel_cols=0, stream=None):
        """Keep a column-panel copy of the entries in the plan (see
        drx_plan_set_panels); later matvecs run one L2-sized panel at a time."""
        check(lib().drx_plan_set_panels(self.plan.handle, self.vcode, _p(self.colind),
                                        _p(self.vals), int(panel_cols), _s(stream)))
        return self.panels()

    def panels(self):
        n, w, s = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        check(lib().drx_plan_panels(self.plan.handle, ctypes.byref(n), ctypes.byref(w),
                                    ctypes.byref(s)))
        return n.value, w.value, s.value

    def update_values(self, vals, stream=None):
        """New values for a plan with a panel layout (re-packed in place)."""
        _vals_ok(vals, self.nnz, self.device)
        check(lib().drx_plan_update_values(self.plan.handle, _p(vals), _s(stream)))
        self.vals = vals

    def transpose(self, stream=None):
        if self._t is None:
            self._t = Transpose(self.plan, self.colind, stream)
        return self._t

    def set_transposed_values(self, vals=None, stream=None):
        """Store `vals` (default: the matrix's) in CSC order in the transpose
        (drx_transpose_set_values): later adjoints with the same, unchanged
        tensor stream them instead of gathering through the permutation."""
        vals = self.vals if vals is None else _vals_ok(vals, self.nnz, self.device)
        t = self.transpose(stream)
        check(lib().drx_transpose_set_values(t.handle, DTYPE_CODE[vals.dtype], _p(vals),
                                             _s(stream)))
        # the tensor itself is kept, so its id cannot be reused by another
        self._tvals_ref = vals
        self._tvals_key = (id(vals), vals._version, vals.dtype)

    def stored_transposed_values(self, vals):
        """True when the transpose holds `vals` (same tensor, unchanged)."""
        return self._tvals_key == (id(vals), vals._version, vals.dtype)

    def rmatvec(self, dy, dx=None, alpha=1.0, beta=0.0, vals=None, stream=None):
        """dx = alpha*A^T*dy + beta*dx."""
        vals = self.vals if vals is None else _vals_ok(vals, self.nnz, self.device)
        k, lddy = _cols(_on(dy, self.device, "dy"), self.shape[0])
        xdt = torch.float64 if vals.dtype == torch.float64 else torch.float32
        if dx is None:
            shape = (self.shape[1],) if dy.dim() == 1 else (k, self.shape[1])
            dx = torch.empty(shape, dtype=xdt, device=dy.device)
        kx, lddx = _cols(_on(dx, self.device, "dx"), self.shape[1])
        if kx != k:
            raise ValueError("dy and dx differ in their number of right-hand sides")
        t = self.transpose(stream)
        vptr = ctypes.c_void_p(0) if self.stored_transposed_values(vals) else _p(vals)
        check(lib().drx_spmv_adjoint(t.handle, DTYPE_CODE[vals.dtype], vptr,
                                     DTYPE_CODE[dy.dtype], _p(dy), lddy,
                                     DTYPE_CODE[dx.dtype], _p(dx), lddx, k,
                                     float(alpha), float(beta), _s(stream)))
        return dx

    def dvals(self, x, dy, out=None, alpha=1.0, beta=0.0, stream=None):
        """out[e] = alpha * sum_j dy[row(e), j] x[col(e), j] + beta*out[e]."""
        k, ldx = _cols(_on(x, self.device, "x"), self.shape[1])
        k2, lddy = _cols(_on(dy, self.device, "dy"), self.shape[0])
        if k != k2:
            raise ValueError("x and dy differ in k")
        odt = torch.float64 if x.dtype == torch.float64 else torch.float32
        if out is None:
            out = torch.empty(self.nnz, dtype=odt, device=x.device)
        _vals_ok(out, self.nnz, self.device)
        check(lib().drx_csr_dvals(self.plan.handle, _p(self.colind),
                                  DTYPE_CODE[x.dtype], _p(x), ldx,
                                  DTYPE_CODE[dy.dtype], _p(dy), lddy, k,
                                  DTYPE_CODE[out.dtype], _p(out), float(alpha),
                                  float(beta), _s(stream)))
        return out

    def fused_residual(self, x, t, y=None, r=None, dvals=None, stream=None):
        """y = A x, r = y - t, dvals[e] = r[row(e)] * x[col(e)] in one pass."""
        ydt = torch.float64 if self.vals.dtype == torch.float64 else torch.float32
        m = self.shape[0]
        if x.dim() != 1 or t.dim() != 1:
            raise ValueError("the fused residual takes one right-hand side (1-D x and t)")
        _cols(_on(x, self.device, "x"), self.shape[1])
        _cols(_on(t, self.device, "t"), m)
        y = torch.empty(m, dtype=ydt, device=x.device) if y is None else y
        r = torch.empty(m, dtype=ydt, device=x.device) if r is None else r
        if dvals is None:
            dvals = torch.empty(self.nnz, dtype=ydt, device=x.device)
        for name, v in (("y", y), ("r", r)):
            _cols(_on(v, self.device, name), m)
            if v.dim() != 1:
                raise ValueError("%s must be 1-D" % name)
        _vals_ok(dvals, self.nnz, self.device)
        check(lib().drx_spmv_fused_residual(
            self.plan.handle, self.vcode, _p(self.colind), _p(self.vals),
            DTYPE_CODE[x.dtype], _p(x), DTYPE_CODE[ydt], _p(t), _p(y), _p(r),
            _p(dvals), _s(stream)))
        return y, r, dvals

    def close(self):
        if self._t is not None:
            self._t.close()
            self._t = None
        self.plan.close()


def _cols(v, n):
    """(k, leading dimension) of a (n,) or (k, n) operand. The kernels read a
    1-D operand as n consecutive elements and a 2-D one as k rows of n
    consecutive elements, so other strides are refused."""
    if v.dim() == 1:
        if v.numel() < n:
            raise ValueError("operand shorter than the matrix dimension")
        if v.numel() > 1 and v.stride(0) != 1:
            raise ValueError("1-D operands must have unit stride (got a strided view; "
                             "pass .contiguous())")
        return 1, max(n, 1)
    if v.dim() == 2:
        if v.shape[1] < n or (v.shape[1] > 1 and v.stride(1) != 1):
            raise ValueError("2-D operands are (k, n) with unit column stride")
        if v.shape[0] > 1 and v.stride(0) < n:
            raise ValueError("2-D operands need a row stride of at least n")
        return v.shape[0], max(v.stride(0), n, 1)
    raise ValueError("operand must be 1-D or 2-D")


def _on(t, device, name):
    """Refuse an operand that is not a tensor on the matrix's device (a host
    pointer handed to a kernel would fault the GPU)."""
    if not isinstance(t, torch.Tensor):
        raise TypeError("%s must be a torch.Tensor" % name)
    if t.device != device:
        raise ValueError("%s is on %s; the matrix is on %s" % (name, t.device, device))
    if t.dtype not in DTYPE_CODE:
        raise ValueError("%s has unsupported dtype %s" % (name, t.dtype))
    return t


def _vals_ok(vals, nnz, device):
    _on(vals, device, "vals")
    if vals.dim() != 1 or vals.numel() != nnz:
        raise ValueError("vals must be 1-D with nnz = %d elements (got shape %s)"
                         % (nnz, tuple(vals.shape)))
    if nnz > 1 and vals.stride(0) != 1:
        raise ValueError("vals must be contiguous")
    return vals


def from_host(rowptr, colind, vals, shape, device="cuda", itype=None,
              vdtype=None, block_nnz=0):
    """Upload host arrays (numpy or torch) as a full CSR matrix."""
    rp = torch.as_tensor(rowptr)
    ci = torch.as_tensor(colind)
    va = torch.as_tensor(vals)
    itype = itype or (torch.int64 if rp.dtype in (torch.int64, torch.uint64)
                      else torch.int32)
    rp = rp.to(torch.int64).to(itype).to(device)
    ci = ci.to(torch.int64).to(itype).to(device)
    va = va.to(vdtype or va.dtype).to(device)
    return CsrMatrix(rp, ci, va, shape, base=0, has_sentinel=True,
                     block_nnz=block_nnz)
