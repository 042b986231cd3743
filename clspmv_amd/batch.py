"""Batched multi-right-hand-side SpMV (include/clspmv_batch.h).

    b = Batch(ctx, [csr0, bsr1, ...], ks=[1, 4, 16, ...])   # CSR or BSR r x r
                                                            # (r = 2..4, k 4 or 16)
    b.set_x(i, X_i)            # X_i: (cols_i, k_i) float32, row-major
    b.capture()                # optional: replay one hipGraph per run
    b.run(1)                   # Y_i += A_i X_i for every i, one launch
    Y_i = b.get_y(i)           # (rows_i, k_i)
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import check, lib


def default_tile(k: int) -> L.SpmmTile:
    return lib().clspmv_spmm_default_tile(k)


class Batch:
    def __init__(self, ctx, mats, ks, tiles=None):
        self.ctx = ctx
        self.mats = list(mats)  # keep the host CSR alive for the call
        n = len(self.mats)
        self.ks = [int(k) for k in ks]
        self.shapes = [(m.s.matinfo.height, m.s.matinfo.width) for m in self.mats]
        ents = (L.BatchEntry * n)()
        for i, m in enumerate(self.mats):
            is_bsr = isinstance(m.s, L.Bsr)
            ents[i] = L.BatchEntry(None if is_bsr else C.cast(C.pointer(m.s), C.c_void_p).value,
                                   C.cast(C.pointer(m.s), C.c_void_p).value if is_bsr else None, self.ks[i])
        tarr = None
        if tiles is not None:
            tarr = (L.SpmmTile * n)()
            for i, t in enumerate(tiles):
                tarr[i] = t if isinstance(t, L.SpmmTile) else L.SpmmTile(*t)
        self.h = C.c_void_p()
        check(lib().clspmv_batch_create_entries(ctx.h, n, ents, tarr, C.byref(self.h)), "batch_create_entries")

    def __len__(self):
        return len(self.ks)

    def set_x(self, i, x):
        rows, cols = self.shapes[i]
        x = np.ascontiguousarray(x, np.float32).reshape(cols, self.ks[i])
        check(lib().clspmv_batch_set_x(self.h, i, x.ctypes.data_as(C.POINTER(C.c_float))), "batch_set_x")

    def set_y(self, i, y):
        rows, cols = self.shapes[i]
        y = np.ascontiguousarray(y, np.float32).reshape(rows, self.ks[i])
        check(lib().clspmv_batch_set_y(self.h, i, y.ctypes.data_as(C.POINTER(C.c_float))), "batch_set_y")

    def get_y(self, i):
        rows, cols = self.shapes[i]
        y = np.empty((rows, self.ks[i]), np.float32)
        check(lib().clspmv_batch_get_y(self.h, i, y.ctypes.data_as(C.POINTER(C.c_float))), "batch_get_y")
        return y

    def run(self, ntimes=1):
        check(lib().clspmv_batch_run(self.h, ntimes), "batch_run")

    def capture(self):
        check(lib().clspmv_batch_capture(self.h), "batch_capture")

    def stats(self):
        a, b, c = C.c_longlong(), C.c_longlong(), C.c_longlong()
        it, g = C.c_int(), C.c_int()
        check(lib().clspmv_batch_stats(self.h, C.byref(a), C.byref(b), C.byref(c), C.byref(it), C.byref(g)), "stats")
        return {"nnz": a.value, "nnz_rhs": b.value, "alg_bytes": c.value, "items": it.value, "grid": g.value}

    def free(self):
        if getattr(self, "h", None) and self.ctx.h:
            lib().clspmv_batch_free(self.h)
        self.h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
