"""The MI355X table profiler: its synthetic matrices are exactly what the
library's conversions make of the same banded COO (so a table record
measures the matrix it names), and the committed MI355X tables load in the
reference's schema and drive the analyzer.  CPU only."""
import ctypes as C
import os

import numpy as np
import pytest

from clspmv_amd import _lib as L
from clspmv_amd import matrix as M
from clspmv_amd import profiler as P

HERE = os.path.dirname(os.path.abspath(__file__))
MI355X = os.path.join(os.path.dirname(HERE), "clspmv_amd", "tables", "mi355x")


def _arr(p, n):
    return np.ctypeslib.as_array(p, shape=(n,)).copy()


def _coo(held):
    row, col, val = held._keep
    return M.CooMatrix.from_arrays(row, col, val, held.s.matinfo.width, held.s.matinfo.height)


@pytest.mark.parametrize("k", [1, 4, 7])
def test_flat_generators_match_conversions(k):
    dim = 1024
    coo = _coo(P.make_coo(dim, k)[0])
    held, nnz = P.make_ell(dim, k)
    e = M.coo2ell(coo)
    assert nnz == coo.nnz and np.array_equal(_arr(e.s.ell_col_id, dim * k), held._keep[0])
    held, _ = P.make_csr(dim, k)
    c = M.coo2csr(coo)
    assert np.array_equal(_arr(c.s.csr_row_ptr, dim + 1), held._keep[0])
    assert np.array_equal(_arr(c.s.csr_col_id, dim * k), held._keep[1])
    for S in (32, 128):
        held, _ = P.make_sell(dim, k, S)
        s = M.coo2sell(coo, S)
        assert np.array_equal(_arr(s.s.sell_slice_ptr, dim // S + 1), held._keep[0])
        assert np.array_equal(_arr(s.s.sell_col_id, dim * k), held._keep[1])


@pytest.mark.parametrize("nd", [1, 5, 9])
def test_dia_generators_match_conversions(nd):
    dim = 1024
    held, nnz = P.make_dia(dim, nd)
    offs = held._keep[0]
    r = np.concatenate([np.arange(max(0, -o), min(dim, dim - o)) for o in offs]).astype(np.int32)
    c = np.concatenate([np.arange(max(0, -o), min(dim, dim - o)) + o for o in offs]).astype(np.int32)
    o = np.lexsort((c, r))
    coo = M.CooMatrix.from_arrays(r[o], c[o], np.ones(r.size, np.float32), dim, dim)
    d = M.coo2dia(coo)
    assert nnz == coo.nnz and d.s.dia_num == nd
    assert np.array_equal(_arr(d.s.dia_data, nd * dim), held._keep[1])
    b = M.coo2bdia(coo)
    hb, _ = P.make_dia(dim, nd, True)
    assert b.s.bdia_band_num == 1 and np.array_equal(_arr(b.s.bdia_bptr, 2), hb._keep[1])


def test_block_generator_is_banded_dense_blocks():
    coo = P.block_coo(1024, 2, 4, 3)
    assert coo.nnz == 512 * 3 * 8 and coo.is_sorted()
    b = M.coo2b4csr(coo, 4, 2)
    assert b.s.b4csr_block_num == 512 * 3


def test_mi355x_tables_drive_the_analyzer():
    if not os.path.exists(os.path.join(MI355X, "dia.ben")):
        pytest.skip("MI355X tables not generated yet")
    b = C.c_void_p()
    L.check(L.lib().clspmv_bench_load(MI355X.encode(), C.byref(b)), "bench_load")
    sizes = (C.c_int * 10)()
    L.lib().clspmv_bench_sizes(b, sizes)
    assert all(n > 0 for n in sizes)
    m = L.MaxflopInfo()
    L.check(L.lib().clspmv_maxflop(b, 2097152, 55742968, C.byref(m)), "maxflop")
    assert m.dia_max_flop > 0 and m.ell_max_flop > 0 and m.csr_max_flop > 0
    coo = M.gen_stencil27(24, 1)
    mat = L.Cocktail()
    L.lib().clspmv_init_cocktail(C.byref(mat))
    c = coo.copy()
    L.check(L.lib().clspmv_analyze_matrix(b, C.byref(c.s), C.byref(mat), 1.0, 1, 0, 0), "analyze")
    total = sum(getattr(mat, n).matinfo.nnz for n in ("bdia", "dia", "sbell", "bell", "bcsr", "sell", "ell", "csr", "coo")
                if getattr(mat, f"ifuse{n}"))
    assert total == coo.nnz
    L.lib().clspmv_free_cocktail(C.byref(mat))
    L.lib().clspmv_bench_free(b)
