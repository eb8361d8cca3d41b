"""ctypes access to oracle/liboracle.so (the CPU checker; test infrastructure).

Every function here is a thin numpy wrapper over oracle/oracle.cpp; the math
lives there, with citations to the reference files it restates.
"""
import ctypes
import json
import os

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(REPO, "oracle", "liboracle.so")
GOLDEN = os.path.join(REPO, "tests", "golden")

_P = ctypes.c_void_p
_Z = ctypes.c_size_t
_I = ctypes.c_int64


def ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def load_fixture(name):
    d = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    return {k: d[k] for k in d.files}


def manifest():
    with open(os.path.join(GOLDEN, "manifest.json")) as f:
        return json.load(f)


class Oracle:
    def __init__(self):
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle/liboracle.so not built (make -C oracle)")
        self.L = ctypes.CDLL(ORACLE_SO)
        L = self.L
        for t in ("f32", "f64"):
            for nm in ("orc_mp_gemv_row_", "orc_mp_gemv_eq_"):
                getattr(L, nm + t).argtypes = [_Z, _Z, _Z, _P, _P, _P, _P, _Z, _Z, _P]
            getattr(L, "orc_mp_gemv_eq_i32_" + t).argtypes = [_Z, _Z, _Z, _P, _P, _P, _P, _Z, _Z, _P]
            getattr(L, "orc_spmv_ref_" + t).argtypes = [_P, _I, _I, _I, ctypes.c_int, _P, _P, _P, _I, _I, _P, _I, _P]
            getattr(L, "orc_spmv_t_ref_" + t).argtypes = [_P, _I, _I, _I, _I, ctypes.c_int, _P, _P, _P, _I, _I, _P, _I, _P]
            getattr(L, "orc_dvals_ref_" + t).argtypes = [_P, _I, _I, _I, ctypes.c_int, _P, _P, _I, _P, _I, _I, _P]
            for nm in ("orc_gemveq_cpu_bench_", "orc_gemveq_cpu_bench_i32_"):
                fn = getattr(L, nm + t)
                fn.argtypes = [_Z, _Z, _Z, _P, _P, _P, _P, _Z, _Z, ctypes.c_int, ctypes.c_int, _P, _P]
                fn.restype = ctypes.c_double
        L.orc_row_partition.argtypes = [_Z, _Z, _P, _Z, _P, _P, _P]
        L.orc_eq_partition.argtypes = [_Z, _Z, _P, _Z, _P, _P, _P, _P]
        L.orc_csr_transpose.argtypes = [_P, _I, _I, _I, _I, ctypes.c_int, _P, _P, _P, _P]
        for nm in ("orc_gen_random_csr_f32_u64", "orc_gen_random_csr_f32_i64"):
            getattr(L, nm).argtypes = [_Z, _Z, ctypes.c_double, ctypes.c_uint, _P, _P, _P]
            getattr(L, nm).restype = _Z
        L.orc_gen_band_csr_f64_i64.argtypes = [_Z, _Z, _Z, _P, _P, _P]
        L.orc_gen_band_csr_f64_i64.restype = _Z
        for t in ("f32", "f64"):
            getattr(L, "orc_gemm_ref_" + t).argtypes = [_I, _I, _I, ctypes.c_double, _P, _I, _P,
                                                        _I, ctypes.c_double, _P, _I, _P, _P]

    # -- partitions ---------------------------------------------------------
    def row_partition(self, rowptr, nnz, P):
        rp = np.ascontiguousarray(rowptr, dtype=np.int64)
        m = rp.size - 1
        seg = np.zeros(1, np.uint64)
        off = np.zeros(P, np.uint64)
        sz = np.zeros(P, np.uint64)
        self.L.orc_row_partition(m, nnz, ptr(rp), P, ptr(seg), ptr(off), ptr(sz))
        return int(seg[0]), off.astype(np.int64), sz.astype(np.int64)

    def eq_partition(self, rowptr, nnz, P):
        rp = np.ascontiguousarray(rowptr, dtype=np.int64)
        m = rp.size - 1
        seg = np.zeros(1, np.uint64)
        tot = np.zeros(1, np.uint64)
        off = np.zeros(P, np.uint64)
        sz = np.zeros(P, np.uint64)
        self.L.orc_eq_partition(m, nnz, ptr(rp), P, ptr(seg), ptr(off), ptr(sz), ptr(tot))
        return int(seg[0]), off.astype(np.int64), sz.astype(np.int64), int(tot[0])

    # -- mp gemv restatement --------------------------------------------------
    def mp_gemv(self, kind, rowptr, colind, vals, x, shape, P):
        """kind 'row' | 'eq'. x is (k, n) or (n,). Returns (k, m) or (m,)."""
        t = "f32" if vals.dtype == np.float32 else "f64"
        rp = np.ascontiguousarray(rowptr, dtype=np.int64)
        ci = np.ascontiguousarray(colind, dtype=np.int64)
        v = np.ascontiguousarray(vals)
        xx = np.ascontiguousarray(x, dtype=v.dtype)
        k = 1 if xx.ndim == 1 else xx.shape[0]
        m, n = shape
        res = np.zeros((k, m), dtype=v.dtype)
        fn = getattr(self.L, "orc_mp_gemv_%s_%s" % (kind, t))
        fn(m, n, ci.size, ptr(rp), ptr(ci), ptr(v), ptr(xx), k, P, ptr(res))
        return res[0] if xx.ndim == 1 else res

    def gemveq_cpu_bench(self, rowptr, colind, vals, x, shape, P, warmup=1, reps=3):
        t = "f32" if vals.dtype == np.float32 else "f64"
        i32 = rowptr.dtype == np.int32
        rp = np.ascontiguousarray(rowptr)
        ci = np.ascontiguousarray(colind)
        v = np.ascontiguousarray(vals)
        xx = np.ascontiguousarray(x, dtype=v.dtype)
        k = 1 if xx.ndim == 1 else xx.shape[0]
        m, n = shape
        res = np.zeros((k, m), dtype=v.dtype)
        times = np.zeros(max(reps, 1))
        fn = getattr(self.L, "orc_gemveq_cpu_bench_%s%s" % ("i32_" if i32 else "", t))
        med = fn(m, n, ci.size, ptr(rp), ptr(ci), ptr(v), ptr(xx), k, P, warmup, reps,
                 ptr(res), ptr(times))
        return (res[0] if xx.ndim == 1 else res), med, times

    # -- extended-precision oracle -----------------------------------------
    def gemm(self, a, b, c=None, alpha=1.0, beta=1.0):
        """(alpha a @ b + beta c, |alpha||a||b| + |beta||c|) as float64 from
        2-D row-major float32/float64 arrays (c None: beta = 0)."""
        t = "f32" if a.dtype == np.float32 else "f64"
        a = np.ascontiguousarray(a)
        b = np.ascontiguousarray(b, dtype=a.dtype)
        m, k = a.shape
        n = b.shape[1]
        if c is None:
            c, beta = np.zeros((m, n), a.dtype), 0.0
        c = np.ascontiguousarray(c, dtype=a.dtype)
        out = np.zeros((m, n))
        absout = np.zeros((m, n))
        getattr(self.L, "orc_gemm_ref_" + t)(m, n, k, alpha, ptr(a), max(k, 1), ptr(b), max(n, 1),
                                             beta, ptr(c), max(n, 1), ptr(out), ptr(absout))
        return out, absout

    def spmv(self, rowptr, colind, vals, x, nrows=None, base=0, nnz=None,
             sentinel=True, ldx=None):
        """(y, |A||x|) as float64; x (n,) or (k, n)."""
        t = "f32" if vals.dtype == np.float32 else "f64"
        rp = np.ascontiguousarray(rowptr, dtype=np.int64)
        ci = np.ascontiguousarray(colind, dtype=np.int64)
        v = np.ascontiguousarray(vals)
        if v.dtype not in (np.float32, np.float64):
            v = v.astype(np.float64)
            t = "f64"
        xx = np.ascontiguousarray(x, dtype=v.dtype)
        k = 1 if xx.ndim == 1 else xx.shape[0]
        ldx = xx.shape[-1] if ldx is None else ldx
        nrows = (rp.size - 1 if sentinel else rp.size) if nrows is None else nrows
        nnz = ci.size if nnz is None else nnz
        y = np.zeros((k, nrows))
        a = np.zeros((k, nrows))
        getattr(self.L, "orc_spmv_ref_" + t)(ptr(rp), base, nrows, nnz, int(sentinel),
                                             ptr(ci), ptr(v), ptr(xx), ldx, k, ptr(y),
                                             nrows, ptr(a))
        return (y[0], a[0]) if xx.ndim == 1 else (y, a)

    def spmv_t(self, rowptr, colind, vals, dy, ncols, base=0, sentinel=True):
        t = "f32" if vals.dtype == np.float32 else "f64"
        rp = np.ascontiguousarray(rowptr, dtype=np.int64)
        ci = np.ascontiguousarray(colind, dtype=np.int64)
        v = np.ascontiguousarray(vals)
        d = np.ascontiguousarray(dy, dtype=v.dtype)
        k = 1 if d.ndim == 1 else d.shape[0]
        nrows = d.shape[-1]
        dx = np.zeros((k, ncols))
        a = np.zeros((k, ncols))
        getattr(self.L, "orc_spmv_t_ref_" + t)(ptr(rp), base, nrows, ncols, ci.size,
                                               int(sentinel), ptr(ci), ptr(v), ptr(d),
                                               nrows, k, ptr(dx), ncols, ptr(a))
        return (dx[0], a[0]) if d.ndim == 1 else (dx, a)

    def dvals(self, rowptr, colind, x, dy, base=0, sentinel=True):
        t = "f32" if x.dtype == np.float32 else "f64"
        rp = np.ascontiguousarray(rowptr, dtype=np.int64)
        ci = np.ascontiguousarray(colind, dtype=np.int64)
        xx = np.ascontiguousarray(x)
        d = np.ascontiguousarray(dy, dtype=xx.dtype)
        k = 1 if xx.ndim == 1 else xx.shape[0]
        nrows = d.shape[-1]
        out = np.zeros(ci.size)
        getattr(self.L, "orc_dvals_ref_" + t)(ptr(rp), base, nrows, ci.size, int(sentinel),
                                              ptr(ci), ptr(xx), xx.shape[-1], ptr(d),
                                              nrows, k, ptr(out))
        return out

    def transpose(self, rowptr, colind, ncols, base=0, sentinel=True):
        rp = np.ascontiguousarray(rowptr, dtype=np.int64)
        ci = np.ascontiguousarray(colind, dtype=np.int64)
        nrows = rp.size - 1 if sentinel else rp.size
        cp = np.zeros(ncols + 1, np.int64)
        ri = np.zeros(ci.size, np.int64)
        pm = np.zeros(ci.size, np.int64)
        self.L.orc_csr_transpose(ptr(rp), base, nrows, ncols, ci.size, int(sentinel),
                                 ptr(ci), ptr(cp), ptr(ri), ptr(pm))
        return cp, ri, pm

    # -- generators -----------------------------------------------------------
    def gen_random_csr(self, m, n, density, seed=0, itype="u64"):
        fn = getattr(self.L, "orc_gen_random_csr_f32_" + itype)
        nnz = fn(m, n, density, seed, None, None, None)
        idt = np.uint64 if itype == "u64" else np.int64
        v = np.zeros(nnz, np.float32)
        rp = np.zeros(m + 1, idt)
        ci = np.zeros(nnz, idt)
        fn(m, n, density, seed, ptr(v), ptr(rp), ptr(ci))
        return rp, ci, v

    def gen_band_csr(self, n, up, down):
        nnz = self.L.orc_gen_band_csr_f64_i64(n, up, down, None, None, None)
        v = np.zeros(nnz, np.float64)
        rp = np.zeros(n + 1, np.int64)
        ci = np.zeros(nnz, np.int64)
        self.L.orc_gen_band_csr_f64_i64(n, up, down, ptr(v), ptr(rp), ptr(ci))
        return rp, ci, v


def fp_equal(a, b, eps):
    """The reference's fp_equal (test/gtest/include/common-tests.hpp:82-95),
    elementwise: equal, or |a-b| < max(min_normal, eps*(|a|+|b|))."""
    a = np.asarray(a)
    b = np.asarray(b)
    if a.shape != b.shape:
        return False
    tiny = np.finfo(a.dtype).tiny
    norm = np.minimum(np.abs(a) + np.abs(b), np.finfo(np.float32).max)
    ok = (a == b) | (np.abs(a - b) < np.maximum(tiny, eps * norm))
    return bool(ok.all())


def componentwise_err(y, ystar, absy):
    """max_i |y_i - y*_i| / (|A||x|)_i (rows with |A||x| = 0 must be exact)."""
    y = np.asarray(y, dtype=np.float64)
    d = np.abs(y - ystar)
    den = np.where(absy > 0, absy, 1.0)
    bad_zero = (absy == 0) & (d != 0)
    if bad_zero.any():
        return np.inf
    return float((d / den).max()) if d.size else 0.0
