"""ctypes binding of libdrx_spmv.so (the C ABI declared in include/drx_spmv.h).

The shared library is built in-tree (``make -C distributed-ranges_amd``) into
``distributed-ranges_amd/lib/libdrx_spmv.so``. Loading fails loudly when it is
missing: there is no fallback implementation.
"""
import ctypes
import os

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libdrx_spmv.so")

DRX_OK = 0
DRX_ERR_INVALID = -1
DRX_ERR_UNSUPPORTED = -2
DRX_ERR_HIP = -3
DRX_ERR_NOMEM = -4

F64, F32, BF16, F16 = 0, 1, 2, 3
I32, I64 = 0, 1

c_i64 = ctypes.c_int64
c_int = ctypes.c_int
c_dbl = ctypes.c_double
c_vp = ctypes.c_void_p
c_u64 = ctypes.c_uint64

# name -> (restype, argtypes)
SIGNATURES = {
    "drx_version": (c_int, []),
    "drx_last_error": (ctypes.c_char_p, []),
    "drx_mm_read_csr": (c_int, [ctypes.c_char_p, c_int, c_int, c_int, c_int,
                                ctypes.POINTER(c_i64), ctypes.POINTER(c_i64),
                                ctypes.POINTER(c_i64), ctypes.POINTER(c_vp),
                                ctypes.POINTER(c_vp), ctypes.POINTER(c_vp)]),
    "drx_host_free": (None, [c_vp]),
    "drx_gather": (c_int, [c_vp, c_i64, c_int, c_vp, c_i64, c_vp, c_vp]),
    "drx_gemm": (c_int, [c_int, c_i64, c_i64, c_i64, ctypes.c_double, c_vp, c_i64, c_vp, c_i64,
                         ctypes.c_double, c_vp, c_i64, c_vp]),
    "drx_tune_set": (c_int, [ctypes.c_char_p, c_i64]),
    "drx_tune_get": (c_int, [ctypes.c_char_p, ctypes.POINTER(c_i64)]),
    "drx_plan_create": (c_int, [ctypes.POINTER(c_vp), c_int, c_vp, c_i64, c_i64,
                                c_i64, c_i64, c_int, c_i64, c_vp]),
    "drx_plan_destroy": (c_int, [c_vp]),
    "drx_plan_info": (c_int, [c_vp, ctypes.POINTER(c_i64), ctypes.POINTER(c_i64)]),
    "drx_plan_set_panels": (c_int, [c_vp, c_int, c_vp, c_vp, c_i64, c_vp]),
    "drx_plan_update_values": (c_int, [c_vp, c_vp, c_vp]),
    "drx_plan_auto_panels": (c_int, [c_vp, c_int, c_vp, c_vp, c_vp, ctypes.POINTER(c_int)]),
    "drx_plan_panels": (c_int, [c_vp, ctypes.POINTER(c_i64), ctypes.POINTER(c_i64),
                                ctypes.POINTER(c_i64)]),
    "drx_spmv": (c_int, [c_vp, c_int, c_vp, c_vp, c_int, c_vp, c_i64, c_int,
                         c_vp, c_i64, c_i64, c_dbl, c_dbl, c_vp]),
    "drx_spmv_csr": (c_int, [c_int, c_vp, c_i64, c_i64, c_i64, c_i64, c_int,
                             c_vp, c_int, c_vp, c_int, c_vp, c_i64, c_int,
                             c_vp, c_i64, c_i64, c_dbl, c_dbl, c_vp]),
    "drx_spmv_csr_noalloc": (c_int, [c_int, c_vp, c_i64, c_i64, c_i64, c_i64, c_int,
                                     c_vp, c_int, c_vp, c_int, c_vp, c_i64, c_int,
                                     c_vp, c_i64, c_i64, c_dbl, c_dbl, c_vp]),
    "drx_spmv_csr_noalloc_max_nnz": (c_i64, []),
    "drx_transpose_create": (c_int, [ctypes.POINTER(c_vp), c_vp, c_vp, c_vp]),
    "drx_transpose_destroy": (c_int, [c_vp]),
    "drx_transpose_arrays": (c_int, [c_vp, ctypes.POINTER(c_vp),
                                     ctypes.POINTER(c_vp), ctypes.POINTER(c_vp)]),
    "drx_transpose_set_values": (c_int, [c_vp, c_int, c_vp, c_vp]),
    "drx_spmv_adjoint": (c_int, [c_vp, c_int, c_vp, c_int, c_vp, c_i64, c_int,
                                 c_vp, c_i64, c_i64, c_dbl, c_dbl, c_vp]),
    "drx_csr_dvals": (c_int, [c_vp, c_vp, c_int, c_vp, c_i64, c_int, c_vp,
                              c_i64, c_i64, c_int, c_vp, c_dbl, c_dbl, c_vp]),
    "drx_spmv_fused_residual": (c_int, [c_vp, c_int, c_vp, c_vp, c_int, c_vp,
                                        c_int, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "drx_gen_uniform": (c_int, [c_int, c_vp, c_i64, c_dbl, c_dbl, c_u64, c_vp]),
    "drx_gen_normal": (c_int, [c_int, c_vp, c_i64, c_u64, c_vp]),
    "drx_gen_zipf_rowptr": (c_int, [c_int, c_vp, c_i64, c_dbl, c_i64, c_u64,
                                    ctypes.POINTER(c_i64), c_vp]),
    "drx_gen_stratified_colind": (c_int, [c_int, c_vp, c_i64, c_i64, c_vp,
                                          c_u64, c_vp]),
    "drx_gen_laplacian_nnz": (c_i64, [c_int, c_i64]),
    "drx_gen_laplacian": (c_int, [c_int, c_i64, c_int, c_vp, c_vp, c_int, c_vp,
                                  c_vp]),
    "drx_gen_band_nnz": (c_i64, [c_i64, c_i64, c_i64]),
    "drx_gen_band": (c_int, [c_i64, c_i64, c_i64, c_int, c_vp, c_vp, c_int,
                             c_vp, c_u64, c_vp]),
    "drx_convert": (c_int, [c_int, c_vp, c_int, c_vp, c_i64, c_vp]),
    "drx_copy_device": (c_int, [c_vp, c_vp, c_i64, c_vp]),
    "drx_stream_read": (c_int, [c_vp, c_i64, c_vp, c_i64, c_vp]),
}


class BatchItem(ctypes.Structure):
    """drx_batch_item (include/drx_spmv.h)."""
    _fields_ = [("plan", c_vp), ("transpose", c_vp), ("colind", c_vp),
                ("vals", c_vp), ("x", c_vp), ("ldx", c_i64), ("y", c_vp),
                ("ldy", c_i64)]


SIGNATURES.update({
    "drx_batch_create": (c_int, [ctypes.POINTER(c_vp), c_i64, ctypes.POINTER(BatchItem),
                                 c_int, c_int, c_int, c_i64, c_dbl, c_dbl, c_vp]),
    "drx_batch_destroy": (c_int, [c_vp]),
    "drx_batch_run": (c_int, [c_vp, c_vp]),
    "drx_batch_graph": (c_int, [c_vp]),
    "drx_batch_graph_launch": (c_int, [c_vp, c_vp]),
    "drx_batch_set_tuning": (c_int, [c_vp, c_i64, c_int, c_int, c_vp]),
    "drx_batch_autotune": (c_int, [c_vp, c_int, ctypes.c_char_p, c_i64, c_vp]),
    "drx_batch_info": (c_int, [c_vp, ctypes.POINTER(c_i64), ctypes.POINTER(c_i64),
                               ctypes.POINTER(c_int), ctypes.POINTER(c_int)]),
    "drx_gen_uniform_rowptr": (c_int, [c_int, c_vp, c_i64, c_i64, c_i64, c_u64,
                                       ctypes.POINTER(c_i64), c_vp]),
})


class DrxError(RuntimeError):
    def __init__(self, status, message):
        super().__init__("drx status %d: %s" % (status, message))
        self.status = status


_lib = None


def lib():
    """Load libdrx_spmv.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                "libdrx_spmv.so not built at %s (run `make -C "
                "distributed-ranges_amd` or __graft_entry__.build())" % LIB_PATH)
        handle = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(status):
    if status != DRX_OK:
        msg = lib().drx_last_error()
        raise DrxError(status, msg.decode() if msg else "")
    return status


def tune_get(key):
    """Current value of a process-wide tuning knob (drx_tune_get)."""
    v = ctypes.c_int64()
    check(lib().drx_tune_get(key.encode(), ctypes.byref(v)))
    return v.value


def tune_set(key, value):
    """Set a process-wide tuning knob (drx_tune_set)."""
    check(lib().drx_tune_set(key.encode(), int(value)))


class tuned:
    """Context manager setting tuning knobs and restoring them on exit:
    ``with drx.tuned(mrhs=0): ...``."""

    def __init__(self, **knobs):
        self.knobs = knobs
        self.saved = {}

    def __enter__(self):
        for k, v in self.knobs.items():
            self.saved[k] = tune_get(k) if k != "block_nnz" else 0
            tune_set(k, v)
        return self

    def __exit__(self, *exc):
        for k, v in self.saved.items():
            tune_set(k, v)
        return False
