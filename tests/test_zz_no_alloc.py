"""drx_spmv_csr_noalloc -- the one-shot product that allocates nothing (its
plan records live in a library-owned device array) -- against drx_spmv_csr,
through the C ABI. Two calls in a row on one stream, the second with a
smaller share, each compared with drx_spmv_csr on the same inputs: bit for
bit (both run the same kernels over the same row blocks) and against the
oracle at the FP64 / FP32 tolerance. Also: k = 3 right-hand sides (column
gathers), accumulation (alpha = beta = 1, the dr::sp tile semantics), a
share without a sentinel, and the size limit (rejected before any device
work)."""
import ctypes

import numpy as np
import pytest
import torch

import drx
from drx._lib import DRX_ERR_INVALID, lib
from drx.core import DTYPE_CODE, ITYPE_CODE, _p, _s
from _oracle import componentwise_err

DEV = "cuda"


def _call(fn, A, x, y, k=1, alpha=1.0, beta=0.0, base=0, sent=1, nrows=None, rowptr=None):
    rp = A.rowptr if rowptr is None else rowptr
    m = A.shape[0] if nrows is None else nrows
    vc = DTYPE_CODE[A.vals.dtype]
    return fn(ITYPE_CODE[rp.dtype], _p(rp), base, m, A.shape[1], A.nnz, sent,
              _p(A.colind), vc, _p(A.vals), DTYPE_CODE[x.dtype], _p(x),
              A.shape[1], DTYPE_CODE[y.dtype], _p(y), m, k, alpha, beta, _s())


def test_noalloc_limit_rejected_without_device_work():
    L = lib()
    nmax = L.drx_spmv_csr_noalloc_max_nnz()
    assert nmax > 100_000_000
    st = L.drx_spmv_csr_noalloc(0, None, 0, 1, 1, nmax + 1, 1, None, 1, None, 1,
                                None, 1, 1, None, 1, 1, 1.0, 0.0, None)
    assert st == DRX_ERR_INVALID
    assert b"scratch" in L.drx_last_error()


@pytest.mark.gpu
@pytest.mark.parametrize("vdt", [torch.float64, torch.float32])
def test_noalloc_matches_one_shot_twice_second_smaller(orc, vdt):
    L = lib()
    tol = 1e-12 if vdt == torch.float64 else 1e-5
    shapes = [(300_000, 250_000, 1, 400, 11), (40_000, 90_000, 1, 3000, 12)]
    for (m, n, lmin, lmax, seed) in shapes:
        A0 = drx.gen.uniform_rows_csr(m, n, lmin, lmax, seed=seed)
        A = drx.CsrMatrix(A0.rowptr, A0.colind, A0.vals.to(vdt), A0.shape)
        x = drx.gen.uniform(n, vdt, -1, 1, seed=seed + 1)
        y_ref = torch.full((m,), 3.0, dtype=vdt, device=DEV)
        y_na = torch.full((m,), -5.0, dtype=vdt, device=DEV)
        print("noalloc case m=%d n=%d nnz=%d %s" % (m, n, A.nnz, vdt), flush=True)
        assert _call(L.drx_spmv_csr, A, x, y_ref) == 0
        assert _call(L.drx_spmv_csr_noalloc, A, x, y_na) == 0, L.drx_last_error()
        torch.cuda.synchronize()
        assert torch.equal(y_na, y_ref)
        rp, ci, v = (A.rowptr.long().cpu().numpy(), A.colind.long().cpu().numpy(),
                     A.vals.cpu().numpy())
        ys, absy = orc.spmv(rp, ci, v, x.cpu().numpy())
        assert componentwise_err(y_na.cpu().numpy(), ys, absy) < tol
        # accumulate (the dr::sp tile call: alpha = beta = 1)
        acc_ref = y_ref.clone()
        acc_na = y_ref.clone()
        assert _call(L.drx_spmv_csr, A, x, acc_ref, beta=1.0) == 0
        assert _call(L.drx_spmv_csr_noalloc, A, x, acc_na, beta=1.0) == 0
        torch.cuda.synchronize()
        assert torch.equal(acc_na, acc_ref)
        A.close()
        A0.close()


@pytest.mark.gpu
def test_noalloc_multi_rhs_and_share_without_sentinel(orc):
    L = lib()
    m, n, k = 20_000, 30_000, 3
    A = drx.gen.uniform_rows_csr(m, n, 1, 200, seed=31)
    X = drx.gen.uniform(k * n, torch.float32, -1, 1, seed=32).reshape(k, n)
    Y_ref = torch.zeros(k, m, device=DEV)
    Y_na = torch.zeros(k, m, device=DEV)
    with drx.tuned(mrhs=0):  # the same column-gather kernels as the no-alloc call
        assert _call(L.drx_spmv_csr, A, X, Y_ref, k=k) == 0
    assert _call(L.drx_spmv_csr_noalloc, A, X, Y_na, k=k) == 0
    torch.cuda.synchronize()
    assert torch.equal(Y_na, Y_ref)
    rp, ci, v = A.rowptr.long().cpu().numpy(), A.colind.long().cpu().numpy(), A.vals.cpu().numpy()
    for j in range(k):
        ys, absy = orc.spmv(rp, ci, v, X[j].cpu().numpy())
        assert componentwise_err(Y_na[j].cpu().numpy(), ys, absy) < 1e-5
    # rows 5000..12000 as a csr_row share: absolute offsets, no sentinel
    r0, r1 = 5000, 12000
    e0, e1 = int(A.rowptr[r0]), int(A.rowptr[r1])
    S = drx.CsrMatrix(A.rowptr[r0:r1].clone(), A.colind[e0:e1].clone(),
                      A.vals[e0:e1].clone(), (r1 - r0, n), base=e0, has_sentinel=False)
    x = X[0].contiguous()
    ys_ref = torch.zeros(r1 - r0, device=DEV)
    ys_na = torch.zeros(r1 - r0, device=DEV)
    assert _call(L.drx_spmv_csr, S, x, ys_ref, base=e0, sent=0) == 0
    assert _call(L.drx_spmv_csr_noalloc, S, x, ys_na, base=e0, sent=0) == 0
    torch.cuda.synchronize()
    assert torch.equal(ys_na, ys_ref)
    ys, absy = orc.spmv(rp[r0:r1 + 1] - e0, ci[e0:e1], v[e0:e1], x.cpu().numpy())
    assert componentwise_err(ys_na.cpu().numpy(), ys, absy) < 1e-5
    S.close()
    A.close()
