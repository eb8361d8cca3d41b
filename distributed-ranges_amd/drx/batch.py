"""Batched products in one persistent launch (drx_batch_*), optionally as a
hipGraph, with the library's autotuner. The JSON of a tuning can be cached
to a file and re-applied."""
import ctypes
import json
import os

from ._lib import BatchItem, check, lib
from .core import DTYPE_CODE, CsrMatrix, _cols, _on, _p, _s


class Batch:
    """forward: [(A, x, y)], adjoint: [(A, dy, dx)] with CsrMatrix A and
    device tensors; x/y are (n,) or (k, n). All items share k, alpha, beta.
    An adjoint item whose matrix holds its current values in CSC order
    (CsrMatrix.set_transposed_values) streams that copy; values changed
    after the batch is made then need set_transposed_values again (the
    batch reads the copy), others read A.vals through the permutation."""

    def __init__(self, forward=(), adjoint=(), k=1, alpha=1.0, beta=0.0, stream=None):
        forward, adjoint = list(forward), list(adjoint)
        n = len(forward) + len(adjoint)
        if n == 0:
            raise ValueError("empty batch")
        items = (BatchItem * n)()
        keep = []
        ref = None
        for i, (A, x, y) in enumerate(forward):
            kx, ldx = _cols(_on(x, A.device, "x"), A.shape[1])
            ky, ldy = _cols(_on(y, A.device, "y"), A.shape[0])
            if kx != k or ky != k:
                raise ValueError("batch item operands must hold k = %d right-hand sides" % k)
            items[i] = BatchItem(A.plan.handle.value, None, A.colind.data_ptr(),
                                 A.vals.data_ptr(), x.data_ptr(), ldx, y.data_ptr(), ldy)
            keep += [A, x, y]
            ref = ref or (A.vals.dtype, x.dtype, y.dtype)
        for j, (A, dy, dx) in enumerate(adjoint):
            t = A.transpose(stream)
            ky, lddy = _cols(_on(dy, A.device, "dy"), A.shape[0])
            kx, lddx = _cols(_on(dx, A.device, "dx"), A.shape[1])
            if kx != k or ky != k:
                raise ValueError("batch item operands must hold k = %d right-hand sides" % k)
            vptr = None if A.stored_transposed_values(A.vals) else A.vals.data_ptr()
            items[len(forward) + j] = BatchItem(None, t.handle.value, None, vptr,
                                                dy.data_ptr(), lddy, dx.data_ptr(), lddx)
            keep += [A, dy, dx]
            ref = ref or (A.vals.dtype, dy.dtype, dx.dtype)
        self._keep = keep
        self.n = n
        h = ctypes.c_void_p()
        check(lib().drx_batch_create(ctypes.byref(h), n, items, DTYPE_CODE[ref[0]],
                                     DTYPE_CODE[ref[1]], DTYPE_CODE[ref[2]], int(k),
                                     float(alpha), float(beta), _s(stream)))
        self.handle = h

    def run(self, stream=None):
        check(lib().drx_batch_run(self.handle, _s(stream)))

    def graph(self):
        check(lib().drx_batch_graph(self.handle))

    def graph_launch(self, stream=None):
        check(lib().drx_batch_graph_launch(self.handle, _s(stream)))

    def set_tuning(self, block_nnz, gshift, wg_per_cu, stream=None):
        check(lib().drx_batch_set_tuning(self.handle, int(block_nnz), int(gshift),
                                         int(wg_per_cu), _s(stream)))

    def autotune(self, iters=3, cache=None, key=None, stream=None):
        """Time every tuning on this batch and keep the fastest. With `cache`
        (a JSON file) and `key`, a stored tuning is applied instead, and a new
        one is stored."""
        if cache and key and os.path.exists(cache):
            with open(cache) as f:
                stored = json.load(f)
            if key in stored:
                t = stored[key]
                self.set_tuning(t["block_nnz"], t["gshift"], t["wg_per_cu"], stream)
                return dict(t, cached=True)
        buf = ctypes.create_string_buffer(256)
        check(lib().drx_batch_autotune(self.handle, int(iters), buf, 256, _s(stream)))
        t = json.loads(buf.value.decode())
        if cache and key:
            stored = {}
            if os.path.exists(cache):
                with open(cache) as f:
                    stored = json.load(f)
            stored[key] = t
            os.makedirs(os.path.dirname(os.path.abspath(cache)), exist_ok=True)
            with open(cache, "w") as f:
                json.dump(stored, f, indent=1, sort_keys=True)
        return t

    def info(self):
        w, b = ctypes.c_int64(), ctypes.c_int64()
        g, c = ctypes.c_int(), ctypes.c_int()
        check(lib().drx_batch_info(self.handle, ctypes.byref(w), ctypes.byref(b),
                                   ctypes.byref(g), ctypes.byref(c)))
        return {"work_units": w.value, "block_nnz": b.value, "gshift": g.value,
                "wg_per_cu": c.value}

    def close(self):
        if getattr(self, "handle", None):
            lib().drx_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
