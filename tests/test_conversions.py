"""Bit-exact parity of libclspmv's host conversions with the reference's
compiled conversions (golden fixtures written by tests/golden/make_golden.py
from oracle/_ref), the reference's worked layout examples
(A/include/matrix_storage.h:169-309), round trips and invariants.
CPU only."""
import os

import numpy as np
import pytest

from clspmv_amd import _lib as L
from clspmv_amd import matrix as M

HERE = os.path.dirname(os.path.abspath(__file__))
G = np.load(os.path.join(HERE, "golden", "conversions.npz"))
CASES = sorted({k.split("/")[0] for k in G.files})
SHAPES = [(bh, bw) for bw in (4, 8) for bh in (1, 2, 4, 8)]


def coo_of(case):
    h, w = (int(v) for v in G[f"{case}/shape"])
    return M.CooMatrix.from_arrays(G[f"{case}/in_row"], G[f"{case}/in_col"], G[f"{case}/in_val"], w, h)


def eq(a, b):
    return a.shape == b.shape and np.array_equal(a.view(np.uint32) if a.dtype == np.float32 else a,
                                                 b.view(np.uint32) if b.dtype == np.float32 else b)


@pytest.mark.parametrize("case", CASES)
def test_csr_and_pad(case):
    csr = M.coo2csr(coo_of(case))
    assert eq(csr.row_ptr, G[f"{case}/csr_row_ptr"])
    for a in (4, 16):
        if f"{case}/padcsr{a}_row_ptr" not in G.files:
            continue
        p = M.pad_csr(csr, a)
        assert eq(p.row_ptr, G[f"{case}/padcsr{a}_row_ptr"])
        assert eq(p.col_id, G[f"{case}/padcsr{a}_col"])
        assert eq(p.data, G[f"{case}/padcsr{a}_val"])


@pytest.mark.parametrize("case", CASES)
def test_dia_bdia(case):
    ok = int(G[f"{case}/dia_ok"][0])
    coo = coo_of(case)
    if ok == -1:
        pytest.skip("no reference fixture (height > width or empty)")
    if ok == 0:
        with pytest.raises(L.ClspmvError) as e:
            M.coo2dia(coo)
        assert e.value.code == L.E_PADDING
        return
    d = M.coo2dia(coo)
    meta = G[f"{case}/dia_meta"]
    assert [d.s.dia_num, d.s.dia_length, d.s.dia_length_aligned] == list(meta)
    assert eq(d.offsets, G[f"{case}/dia_offsets"])
    assert eq(d.data, G[f"{case}/dia_data"])
    b = M.coo2bdia(coo_of(case))
    assert eq(b.offsets, G[f"{case}/bdia_offsets"])
    assert eq(b.bptr, G[f"{case}/bdia_bptr"])
    assert eq(b.data, G[f"{case}/dia_data"])


@pytest.mark.parametrize("case", CASES)
def test_ell(case):
    e = M.coo2ell(coo_of(case))
    assert [e.s.ell_num, e.s.ell_height_aligned] == list(G[f"{case}/ell_meta"])
    assert eq(e.col_id, G[f"{case}/ell_col"])
    assert eq(e.data, G[f"{case}/ell_val"])


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("S", [32, 128])
def test_sell(case, S):
    s = M.coo2sell(coo_of(case), S)
    assert eq(s.slice_ptr, G[f"{case}/sell{S}_ptr"])
    assert eq(s.col_id, G[f"{case}/sell{S}_col"])
    assert eq(s.data, G[f"{case}/sell{S}_val"])


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("bh,bw", SHAPES)
def test_block_formats(case, bh, bw):
    tag = f"{bh}x{bw}"
    b = M.coo2b4csr(coo_of(case), bw, bh)
    assert eq(b.row_ptr, G[f"{case}/b4csr{tag}_row_ptr"])
    assert eq(b.col_id, G[f"{case}/b4csr{tag}_col"])
    assert eq(b.data, G[f"{case}/b4csr{tag}_val"])
    e = M.coo2b4ell(coo_of(case), bw, bh)
    brows, Hc, K, F = (int(v) for v in G[f"{case}/b4ell{tag}_meta"])
    assert [e.s.b4ell_row_num, e.s.b4ell_height_aligned, e.s.b4ell_block_num, e.s.b4ell_float4_aligned] == \
        [brows, Hc, K, F]
    assert eq(np.ascontiguousarray(e.col_id.reshape(K, Hc)[:, :brows]), G[f"{case}/b4ell{tag}_col"])
    assert eq(np.ascontiguousarray(e.data.reshape(K * (bw // 4) * bh, F)[:, :4 * brows]), G[f"{case}/b4ell{tag}_val"])
    s = M.coo2sbell(coo_of(case), bw, bh, 32)
    assert eq(s.slice_ptr, G[f"{case}/sbell{tag}_ptr"])
    assert eq(s.col_id, G[f"{case}/sbell{tag}_col"])
    assert eq(s.data, G[f"{case}/sbell{tag}_val"])


# ---- worked examples of A/include/matrix_storage.h (values are the letters'
# positions: 0..f, g..v) --------------------------------------------------

def _example_matrix():
    # |0 1 2 3| g h i j|
    # |4 5 6 7| k l m n|
    # |8 9 a b| o p q r|
    # |c d e f| s t u v|     4 x 8, value = 1 + position code
    code = {}
    for i, ch in enumerate("0123456789abcdef"):
        code[ch] = i
    for i, ch in enumerate("ghijklmnopqrstuv"):
        code[ch] = 16 + i
    rows = ["0123ghij", "4567klmn", "89abopqr", "cdefstuv"]
    r, c, v = [], [], []
    for i, s in enumerate(rows):
        for j, ch in enumerate(s):
            r.append(i)
            c.append(j)
            v.append(1.0 + code[ch])
    return M.CooMatrix.from_arrays(np.array(r), np.array(c), np.array(v, np.float32), 8, 4), code


def test_b4csr_worked_example_2x4_2x8():
    coo, code = _example_matrix()
    val = lambda s: [1.0 + code[ch] for ch in s]  # noqa: E731
    b = M.coo2b4csr(coo, 4, 2)  # 2 x 4 blocks (matrix_storage.h:248-252)
    assert list(b.row_ptr) == [0, 2, 4]
    assert list(b.col_id) == [0, 1, 0, 1]
    N = b.s.b4csr_aligned_size
    d = b.data.reshape(2, N)  # h = 0, 1 (w4 = 0)
    assert list(d[0, :16]) == val("0123ghij89abopqr")
    assert list(d[1, :16]) == val("4567klmncdefstuv")
    b = M.coo2b4csr(coo.copy(), 8, 2)  # 2 x 8 blocks (:255-261)
    assert list(b.row_ptr) == [0, 1, 2]
    assert list(b.col_id) == [0, 0]
    N = b.s.b4csr_aligned_size
    d = b.data.reshape(2, 2, N)  # [w4][h]
    assert list(d[0, 0, :8]) == val("012389ab")
    assert list(d[0, 1, :8]) == val("4567cdef")
    assert list(d[1, 0, :8]) == val("ghijopqr")
    assert list(d[1, 1, :8]) == val("klmnstuv")


def test_b4ell_worked_example_2x4():
    coo, code = _example_matrix()
    val = lambda s: [1.0 + code[ch] for ch in s]  # noqa: E731
    e = M.coo2b4ell(coo, 4, 2)  # (:299-302)
    Hc, F = e.s.b4ell_height_aligned, e.s.b4ell_float4_aligned
    col = e.col_id.reshape(2, Hc)
    assert [col[0, 0], col[0, 1], col[1, 0], col[1, 1]] == [0, 0, 1, 1]
    d = e.data.reshape(2, 2, F)  # [slot][h] for w4 = 0
    assert list(d[0, 0, :8]) == val("012389ab")
    assert list(d[0, 1, :8]) == val("4567cdef")
    assert list(d[1, 0, :8]) == val("ghijopqr")
    assert list(d[1, 1, :8]) == val("klmnstuv")


def test_round_trips_and_invariants():
    rng = np.random.default_rng(5)
    for case in CASES:
        coo = coo_of(case)
        csr = M.coo2csr(coo)
        rows = np.repeat(np.arange(coo.height, dtype=np.int32), np.diff(csr.row_ptr))
        assert np.array_equal(rows, coo.row_id) and np.array_equal(csr.col_id, coo.col_id)
        e = M.coo2ell(coo_of(case))
        H = e.s.ell_height_aligned
        vals = e.data.reshape(-1, H)[:, :coo.height]
        assert np.count_nonzero(vals) == np.count_nonzero(coo.data)
    # exact ELL for a fixed-row-length matrix: no padding at all
    coo = M.gen_uniform_rows(256, 7, 3)
    e = M.coo2ell(coo)
    assert e.s.ell_num == 7 and e.s.ell_height_aligned == 256
    assert np.count_nonzero(e.data) == coo.nnz
    del rng


def test_stencil_dia_closed_form():
    # 27-point stencil on n^3: offsets dz*n^2 + dy*n + dx, 9 bands of width 3
    n = 6
    coo = M.gen_stencil27(n, 1)
    d = M.coo2dia(coo)
    want = sorted(dz * n * n + dy * n + dx for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1))
    assert list(d.offsets) == want
    b = M.coo2bdia(M.gen_stencil27(n, 1))
    assert b.s.bdia_band_num == 9 and list(np.diff(b.bptr)) == [3] * 9
    assert list(b.offsets) == [dz * n * n + dy * n - 1 for dz in (-1, 0, 1) for dy in (-1, 0, 1)]
    assert coo.nnz == (3 * n - 2) ** 3


def test_sort_coo_matches_golden_order():
    # unsorted input: the restated quicksort must give row-major order
    rng = np.random.default_rng(3)
    case = CASES[1]
    coo = coo_of(case)
    perm = rng.permutation(coo.nnz)
    sh = M.CooMatrix.from_arrays(coo.row_id[perm], coo.col_id[perm], coo.data[perm], coo.width, coo.height)
    assert not sh.is_sorted()
    sh.sort()
    assert np.array_equal(sh.row_id, coo.row_id) and np.array_equal(sh.col_id, coo.col_id)
    assert np.array_equal(sh.data, coo.data)
