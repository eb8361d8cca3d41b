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
