"""The C ABI library loads and exports every symbol include/*.h declares
(CPU only: no compute calls; only argument checks that fail before any HIP
call)."""
import ctypes
import glob
import os
import re

import pytest

import drx

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(REPO, "include", "*.h")):
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        for m in re.finditer(r"\b(drx_\w+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = drx.lib()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_bindings_cover_every_declared_symbol():
    from drx._lib import SIGNATURES
    assert set(declared_functions()) <= set(SIGNATURES)


def test_version_and_error_string():
    assert drx.lib().drx_version() >= 1
    assert isinstance(drx.lib().drx_last_error(), bytes)


def test_argument_validation_without_gpu():
    L = drx.lib()
    # null out-pointer, bad itype, negative size: rejected before any HIP call
    assert L.drx_plan_create(None, 0, None, 0, 0, 0, 0, 1, 0, None) == drx._lib.DRX_ERR_INVALID
    h = ctypes.c_void_p()
    assert L.drx_plan_create(ctypes.byref(h), 7, None, 0, 1, 1, 1, 1, 0, None) == drx._lib.DRX_ERR_INVALID
    assert b"itype" in L.drx_last_error()
    assert L.drx_plan_create(ctypes.byref(h), 0, None, 0, -1, 1, 1, 1, 0, None) == drx._lib.DRX_ERR_INVALID
    assert L.drx_plan_create(ctypes.byref(h), 0, None, 0, 4, 4, 4, 1, 0, None) == drx._lib.DRX_ERR_INVALID
    assert L.drx_plan_create(ctypes.byref(h), 0, ctypes.c_void_p(8), 0, 4, 4, 4, 1, 100, None) == drx._lib.DRX_ERR_INVALID
    assert b"block_nnz" in L.drx_last_error()
    assert L.drx_spmv(None, 0, None, None, 0, None, 0, 0, None, 0, 1, 1.0, 0.0, None) == drx._lib.DRX_ERR_INVALID
    assert L.drx_tune_set(b"no_such_knob", 1) == drx._lib.DRX_ERR_INVALID
    assert L.drx_tune_set(b"block_nnz", 100) == drx._lib.DRX_ERR_INVALID
    v = ctypes.c_int64()
    assert L.drx_tune_set(b"block_nnz", 2048) == 0
    assert L.drx_tune_get(b"block_nnz", ctypes.byref(v)) == 0 and v.value == 2048
    assert L.drx_tune_set(b"block_nnz", 0) == 0
    assert drx.tune_get("mrhs") == 1 and drx.tune_get("mrhs_gshift") == 1
    assert drx.tune_get("vec") == 0 and drx.tune_get("grid_cap") == 0
    assert L.drx_tune_set(b"vec", 3) == drx._lib.DRX_ERR_INVALID
    assert drx.tune_get("mrhs_kp_max") == 0 and drx.tune_get("wide") == 1
    assert L.drx_tune_set(b"mrhs_kp_max", 4) == drx._lib.DRX_ERR_INVALID
    assert L.drx_tune_set(b"mrhs", 3) == drx._lib.DRX_ERR_INVALID
    assert L.drx_tune_set(b"mrhs_gshift", 4) == drx._lib.DRX_ERR_INVALID
    with drx.tuned(mrhs=0, mrhs_gshift=3):
        assert drx.tune_get("mrhs") == 0 and drx.tune_get("mrhs_gshift") == 3
    assert drx.tune_get("mrhs") == 1 and drx.tune_get("mrhs_gshift") == 1
    assert L.drx_gen_laplacian_nnz(2, 1000) == 4996000
    assert L.drx_gen_laplacian_nnz(3, 256) == 117047296
    assert L.drx_gen_band_nnz(4000000, 16, 16) == 131999728
    assert L.drx_gen_band_nnz(866, 17, 17) == 30004


def test_gemm_argument_validation_without_gpu():
    L = drx.lib()
    bad = drx._lib.DRX_ERR_INVALID
    p = ctypes.c_void_p(256)
    # lda < k, ldb < n, ldc < n, unsupported vtype, negative size, null operand
    assert L.drx_gemm(drx.F32, 4, 6, 5, 1.0, p, 4, p, 6, 1.0, p, 6, None) == bad
    assert L.drx_gemm(drx.F32, 4, 6, 5, 1.0, p, 5, p, 5, 1.0, p, 6, None) == bad
    assert L.drx_gemm(drx.F64, 4, 6, 5, 1.0, p, 5, p, 6, 1.0, p, 5, None) == bad
    assert L.drx_gemm(drx.BF16, 4, 6, 5, 1.0, p, 5, p, 6, 1.0, p, 6, None) == bad
    assert b"vtype" in L.drx_last_error()
    assert L.drx_gemm(drx.F32, -1, 6, 5, 1.0, p, 5, p, 6, 1.0, p, 6, None) == bad
    assert L.drx_gemm(drx.F32, 4, 6, 5, 1.0, None, 5, p, 6, 1.0, p, 6, None) == bad
    # empty products return before any device work
    assert L.drx_gemm(drx.F32, 0, 6, 5, 1.0, None, 5, None, 6, 1.0, None, 6, None) == 0
    v = ctypes.c_int64()
    assert L.drx_tune_set(b"gemm_tile", 8) == bad
    assert L.drx_tune_set(b"gemm_tile", 2) == 0
    assert L.drx_tune_get(b"gemm_tile", ctypes.byref(v)) == 0 and v.value == 2
    assert L.drx_tune_set(b"gemm_tile", 0) == 0
    assert L.drx_tune_get(b"xstage", ctypes.byref(v)) == 0 and v.value == 1
    assert L.drx_tune_set(b"xstage", 3) == bad
    assert L.drx_tune_set(b"xstage", 2) == 0 and L.drx_tune_set(b"xstage", 1) == 0
    assert L.drx_tune_get(b"rowgather", ctypes.byref(v)) == 0 and v.value == 1
    assert L.drx_tune_set(b"rowgather", 4) == bad
    assert L.drx_tune_set(b"rowgather", 2) == 0 and L.drx_tune_set(b"rowgather", 1) == 0


def test_shim_refuses_bad_operands_without_gpu():
    """The ctypes shim refuses operands the kernels would misread: tensors on
    another device (a host pointer would fault the GPU), strided 1-D views
    (read as if contiguous), values of the wrong length, mismatched k. All
    refusals happen before any library call (meta tensors stand in for the
    matrix's device)."""
    import torch
    from drx.core import CsrMatrix
    dev = torch.device("meta")
    A = CsrMatrix.__new__(CsrMatrix)
    A.shape, A.nnz, A.device = (4, 5), 3, dev
    A.vals = torch.empty(3, device=dev)
    A.colind = torch.empty(3, dtype=torch.int32, device=dev)
    A._t, A._tvals_key = None, None
    x = torch.empty(5, device=dev)
    with pytest.raises(ValueError, match="is on cpu"):
        A.matvec(torch.ones(5))
    with pytest.raises(ValueError, match="y is on cpu"):
        A.matvec(x, y=torch.ones(4))
    X = torch.empty(5, 2, device=dev)
    with pytest.raises(ValueError, match="unit stride"):
        A.matvec(X[:, 0])
    with pytest.raises(ValueError, match="nnz"):
        A.matvec(x, vals=torch.empty(4, device=dev))
    with pytest.raises(ValueError, match="contiguous"):
        A.matvec(x, vals=torch.empty(6, device=dev)[::2])
    with pytest.raises(ValueError, match="right-hand sides"):
        A.matvec(torch.empty(2, 5, device=dev), y=torch.empty(3, 4, device=dev))
    with pytest.raises(ValueError, match="unit column stride"):
        A.matvec(torch.empty(5, 2, device=dev).t())
    with pytest.raises(ValueError, match="dy is on cpu"):
        A.rmatvec(torch.ones(4))
    with pytest.raises(ValueError, match="unit stride"):
        A.rmatvec(torch.empty(4, 2, device=dev)[:, 1])
    with pytest.raises(ValueError, match="x is on cpu"):
        A.dvals(torch.ones(5), torch.empty(4, device=dev))
    with pytest.raises(ValueError, match="nnz"):
        A.dvals(x, torch.empty(4, device=dev), out=torch.empty(2, device=dev))
    with pytest.raises(ValueError, match="t is on cpu"):
        A.fused_residual(x, torch.ones(4))
    with pytest.raises(ValueError, match="nnz"):
        A.update_values(torch.empty(2, device=dev))
    with pytest.raises(ValueError, match="unsupported dtype"):
        A.matvec(torch.empty(5, dtype=torch.int32, device=dev))
