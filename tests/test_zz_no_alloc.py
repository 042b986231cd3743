"""The allocation-free COO entry point (clspmv_coo_spmv_noalloc) against the
COO part it shares its kernels with (clspmv_part_coo method 0 +
clspmv_part_spmv): bitwise equal results on two calls in a row, the second
with a smaller matrix (the shared scratch keeps the first call's carries),
both at oracle parity; sizes beyond its scratch are refused."""
import ctypes as C

import numpy as np
import pytest

from clspmv_amd import _lib as L
from clspmv_amd import matrix as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _noalloc(ctx, coo, x, y0):
    dr = M.DeviceArray.from_numpy(ctx, np.ascontiguousarray(coo.row_id, np.int32))
    dc = M.DeviceArray.from_numpy(ctx, np.ascontiguousarray(coo.col_id, np.int32))
    dv = M.DeviceArray.from_numpy(ctx, np.ascontiguousarray(coo.data, np.float32))
    dx = M.DeviceArray.from_numpy(ctx, x)
    dy = M.DeviceArray.from_numpy(ctx, y0)
    L.check(L.lib().clspmv_coo_spmv_noalloc(ctx.h, coo.height, coo.nnz, dr.ptr, dc.ptr, dv.ptr, dx.ptr, dy.ptr),
            "coo_spmv_noalloc")
    y = dy.download()
    for a in (dr, dc, dv, dx, dy):
        a.free()
    return y


def _part(ctx, coo, x, y0):
    part = M.Part(ctx, L.FMT_COO, coo, 0)
    dx = M.DeviceArray.from_numpy(ctx, x)
    dy = M.DeviceArray.from_numpy(ctx, y0)
    part.spmv(dx, dy, 1)
    y = dy.download()
    part.free()
    dx.free()
    dy.free()
    return y


def test_noalloc_matches_coo_part_twice_shrinking(ctx):
    # a first call with a long row (carries across many chunks) and a ragged
    # end, then a smaller matrix whose carries overwrite part of the first's
    first = M.gen_zipf_rows(60000, 1.8, 16.0, 20000, 51)
    second = M.gen_uniform_rows(3001, 7, 52)
    assert first.nnz % 512 and second.nnz % 512 and second.nnz < first.nnz
    for coo, seed in ((first, 53), (second, 55)):
        x = M.gen_vector(coo.width, seed)
        y0 = M.gen_vector(coo.height, seed + 1)
        ya = _noalloc(ctx, coo, x, y0)
        yb = _part(ctx, coo, x, y0)
        assert np.array_equal(ya.view(np.uint32), yb.view(np.uint32)), coo.nnz
        rep = O.parity_report(ya, coo, x, y0)
        assert rep["ok"], rep


def test_noalloc_refuses_beyond_its_scratch(ctx):
    d = M.DeviceArray(ctx, 64)
    rc = L.lib().clspmv_coo_spmv_noalloc(ctx.h, 10, 167772161, d.ptr, d.ptr, d.ptr, d.ptr, d.ptr)
    assert rc == L.E_TOO_LARGE
    assert b"above" in L.lib().clspmv_last_error()
    # a misaligned array is refused, not read
    rc = L.lib().clspmv_coo_spmv_noalloc(ctx.h, 10, 5, C.c_void_p(d.ptr.value + 4), d.ptr, d.ptr, d.ptr, d.ptr)
    assert rc == L.E_INVALID
    d.free()
