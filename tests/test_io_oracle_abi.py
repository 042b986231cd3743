"""CPU tests: Matrix Market reader parity, the oracle's pins, the C ABI
surface and the host-forced failure paths."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from clspmv_amd import _lib as L
from clspmv_amd import matrix as M
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
MMF = np.load(os.path.join(HERE, "golden", "mmf.npz"))
CONV = np.load(os.path.join(HERE, "golden", "conversions.npz"))


# ------------------------------------------------------------- ReadMMF
@pytest.mark.parametrize("fname", sorted({k.split("/")[0] for k in MMF.files}))
def test_read_mmf_matches_reference(fname):
    """clspmv_read_mmf == the reference's ReadMMF (A/src/fileio.cpp:10-109),
    quirks included (pattern -> 1.0, 'integer' mis-parse, symmetric mirror,
    duplicates kept, quicksort order)."""
    coo = M.CooMatrix.read_mmf(os.path.join(HERE, "golden", "mmf", fname))
    h, w, n = (int(v) for v in MMF[f"{fname}/shape"])
    assert (coo.height, coo.width, coo.nnz) == (h, w, n)
    assert np.array_equal(coo.row_id, MMF[f"{fname}/row"])
    assert np.array_equal(coo.col_id, MMF[f"{fname}/col"])
    assert np.array_equal(coo.data.view(np.uint32), MMF[f"{fname}/val"].view(np.uint32))


def test_mmf_write_read_round_trip(tmp_path):
    coo = M.gen_zipf_rows(500, 1.8, 6.0, 100, 4)
    p = tmp_path / "m.mtx"
    coo.write_mmf(p)
    back = M.CooMatrix.read_mmf(p)
    assert np.array_equal(back.row_id, coo.row_id) and np.array_equal(back.col_id, coo.col_id)
    assert np.array_equal(back.data, coo.data)  # %.9g round-trips float32 exactly


def test_read_mmf_missing_file():
    with pytest.raises(L.ClspmvError) as e:
        M.CooMatrix.read_mmf("/nonexistent/file.mtx")
    assert e.value.code == L.E_IO


def test_binary_coo_round_trip(tmp_path):
    coo = M.gen_uniform_rows(300, 5, 9)
    p = str(tmp_path / "m.bin").encode()
    L.check(L.lib().clspmv_write_coo_bin(p, C.byref(coo.s)), "write")
    back = M.CooMatrix()
    L.check(L.lib().clspmv_read_coo_bin(p, C.byref(back.s)), "read")
    assert np.array_equal(back.row_id, coo.row_id) and np.array_equal(back.data, coo.data)


# ------------------------------------------------------------- oracle pins
@pytest.mark.parametrize("case", sorted({k.split("/")[0] for k in CONV.files}))
def test_oracle_spmv_only_bit_exact_vs_reference(case):
    """The restated spmv_only equals the reference's compiled spmv_only
    bit for bit (same FP32 operation order, no FMA)."""
    h, w = (int(v) for v in CONV[f"{case}/shape"])
    y = O.spmv_only(CONV[f"{case}/in_row"], CONV[f"{case}/in_col"], CONV[f"{case}/in_val"],
                    CONV[f"{case}/spmv_x"], h)
    assert np.array_equal(y.view(np.uint32), CONV[f"{case}/spmv_y"].view(np.uint32))


def test_oracle_closed_forms():
    # 5-point Laplacian, x = 1: interior rows 0, edge rows 1, corners 2 (exact)
    n = 64
    coo = M.gen_laplacian2d(n, n)
    y = O.spmv_only(coo.row_id, coo.col_id, coo.data, np.ones(n * n, np.float32), n * n)
    ix, iy = np.arange(n * n) % n, np.arange(n * n) // n
    edges = (ix == 0).astype(int) + (ix == n - 1) + (iy == 0) + (iy == n - 1)
    assert np.array_equal(y, edges.astype(np.float32))
    # unit-valued 27-point stencil, x = 1: y = number of in-grid neighbours
    n = 9
    st = M.gen_stencil27(n, 3)
    y = O.spmv_only(st.row_id, st.col_id, np.ones(st.nnz, np.float32), np.ones(n ** 3, np.float32), n ** 3)
    cnt = lambda i: 1 + (i > 0) + (i < n - 1)  # noqa: E731
    r = np.arange(n ** 3)
    want = np.array([cnt(a) * cnt(b) * cnt(c) for a, b, c in zip(r % n, (r // n) % n, r // (n * n))], np.float32)
    assert np.array_equal(y, want)


def test_oracle_dense_brute_force():
    rng = np.random.default_rng(11)
    h, w = 37, 29
    A = np.where(rng.random((h, w)) < 0.3, rng.standard_normal((h, w)), 0.0).astype(np.float32)
    r, c = np.nonzero(A)
    x = rng.uniform(-1, 1, w).astype(np.float32)
    y64 = O.spmv_f64(r.astype(np.int32), c.astype(np.int32), A[r, c], x, h)
    assert np.allclose(y64, A.astype(np.float64) @ x.astype(np.float64), rtol=0, atol=1e-12)
    y32 = O.spmv_only(r.astype(np.int32), c.astype(np.int32), A[r, c], x, h)
    assert np.allclose(y32, A.astype(np.float64) @ x, rtol=1e-5, atol=1e-6)
    # a transposed operand is caught
    yt = O.spmv_f64(c.astype(np.int32), r.astype(np.int32), A[r, c], rng.uniform(-1, 1, h).astype(np.float32), w)
    assert yt.shape == (w,) and not np.allclose(yt[:h] if w >= h else yt, y64[:len(yt)])


def test_per_format_oracles_match_dense():
    """The per-format CPU SpMVs (layout indexing) reproduce A x on the
    reference's own layouts of a ragged matrix."""
    case = "rand8_200x180"
    h, w = (int(v) for v in CONV[f"{case}/shape"])
    coo = M.CooMatrix.from_arrays(CONV[f"{case}/in_row"], CONV[f"{case}/in_col"], CONV[f"{case}/in_val"], w, h)
    x = np.random.default_rng(1).uniform(-1, 1, w).astype(np.float32)
    ref = O.spmv_f64(coo.row_id, coo.col_id, coo.data, x, h)
    o = O.lib()
    P = lambda a, t=C.c_int: np.ascontiguousarray(a).ctypes.data_as(C.POINTER(t))  # noqa: E731
    F = C.c_float
    y = np.zeros(h, np.float32)
    e = M.coo2ell(coo)
    o.oracle_ell_spmv(e.s.ell_num, e.s.ell_height_aligned, P(e.col_id), P(e.data, F), P(x, F), h, P(y, F))
    assert np.allclose(y, ref, atol=1e-4)
    y[:] = 0
    s = M.coo2sell(coo, 32)
    o.oracle_sell_spmv(32, s.s.sell_slice_num, P(s.slice_ptr), P(s.col_id), P(s.data, F), P(x, F), h, P(y, F))
    assert np.allclose(y, ref, atol=1e-4)
    for bh, bw in [(2, 4), (8, 8), (1, 8)]:
        y[:] = 0
        b = M.coo2b4csr(coo, bw, bh)
        o.oracle_b4csr_spmv(b.s.b4csr_row_num, bh, bw, b.s.b4csr_aligned_size, P(b.row_ptr), P(b.col_id),
                            P(b.data, F), P(x, F), w, h, P(y, F))
        assert np.allclose(y, ref, atol=1e-4)
        y[:] = 0
        be = M.coo2b4ell(coo, bw, bh)
        o.oracle_b4ell_spmv(be.s.b4ell_row_num, bh, bw, be.s.b4ell_block_num, be.s.b4ell_height_aligned,
                            be.s.b4ell_float4_aligned, P(be.col_id), P(be.data, F), P(x, F), w, h, P(y, F))
        assert np.allclose(y, ref, atol=1e-4)
        y[:] = 0
        sb = M.coo2sbell(coo, bw, bh, 32)
        o.oracle_sbell_spmv(32, sb.s.sbell_slice_num, bh, bw, P(sb.slice_ptr), P(sb.col_id), P(sb.data, F),
                            P(x, F), w, h, P(y, F))
        assert np.allclose(y, ref, atol=1e-4)
    y[:] = 0
    bs = M.coo2bsr(coo, 3, 3)
    o.oracle_bsr_spmv(bs.s.bsr_row_num, 3, 3, bs.s.bsr_block_stride, P(bs.row_ptr), P(bs.col_id), P(bs.data, F),
                      P(x, F), w, h, P(y, F))
    assert np.allclose(y, ref, atol=1e-4)


def _dense(case):
    h, w = (int(v) for v in CONV[f"{case}/shape"])
    A = np.zeros((h, w), np.float64)
    np.add.at(A, (CONV[f"{case}/in_row"], CONV[f"{case}/in_col"]), CONV[f"{case}/in_val"].astype(np.float64))
    return A


def _dia_oracles(offsets, data, L_, bptr, boffsets, x, y0, w, h):
    """oracle_dia_spmv and oracle_bdia_spmv (both y += A x) on one layout."""
    o = O.lib()
    P = lambda a, t=C.c_int: np.ascontiguousarray(a).ctypes.data_as(C.POINTER(t))  # noqa: E731
    F = C.c_float
    yd = np.array(y0, np.float32)
    o.oracle_dia_spmv(len(offsets), L_, P(offsets), P(data, F), P(x, F), w, h, P(yd, F))
    yb = np.array(y0, np.float32)
    o.oracle_bdia_spmv(len(boffsets), L_, P(bptr), P(boffsets), P(data, F), P(x, F), w, h, P(yb, F))
    return yd, yb


def test_dia_bdia_oracles_on_reference_layouts():
    """oracle_dia_spmv / oracle_bdia_spmv (A/include/spmv_serial.h:11-34,
    :59-90) on the DIA and BDIA arrays the reference's own coo2dia /
    coo2bdia wrote (golden lap2d_12): the 5-point Laplacian's closed form
    with x = 1 (interior 0, edge 1, corner 2, exactly), and y0 + A x against
    the dense FP64 product for random x and y0."""
    case = "lap2d_12"
    h, w = (int(v) for v in CONV[f"{case}/shape"])
    _, _, L_ = (int(v) for v in CONV[f"{case}/dia_meta"])
    offs, data = CONV[f"{case}/dia_offsets"], CONV[f"{case}/dia_data"]
    bptr, boffs = CONV[f"{case}/bdia_bptr"], CONV[f"{case}/bdia_offsets"]
    yd, yb = _dia_oracles(offs, data, L_, bptr, boffs, np.ones(w, np.float32), np.zeros(h, np.float32), w, h)
    g = int(round(h ** 0.5))
    i, j = np.divmod(np.arange(h), g)
    expect = (i == 0).astype(np.float32) + (i == g - 1) + (j == 0) + (j == g - 1)
    assert np.array_equal(yd, expect) and np.array_equal(yb, expect)
    rng = np.random.default_rng(3)
    x = rng.uniform(-1, 1, w).astype(np.float32)
    y0 = rng.uniform(-1, 1, h).astype(np.float32)
    yd, yb = _dia_oracles(offs, data, L_, bptr, boffs, x, y0, w, h)
    ref = y0.astype(np.float64) + _dense(case) @ x.astype(np.float64)
    assert np.allclose(yd, ref, rtol=0, atol=1e-5) and np.allclose(yb, ref, rtol=0, atol=1e-5)


@pytest.mark.parametrize("h,w", [(50, 64), (64, 50), (97, 97)])
def test_dia_bdia_oracles_nonsymmetric(h, w):
    """The same oracles on non-symmetric, square and rectangular banded
    matrices (diagonals -7, -2, -1, 0, 3, 4, 5 with random gaps) laid out by
    the library's coo2dia / coo2bdia (their arrays are pinned to the
    reference's by the golden conversion tests): a swapped offset sign
    (A^T x) or a row/column slip fails the comparison, and integer values
    with x in {-2..2} make the dense product exact."""
    rng = np.random.default_rng(h * 1000 + w)
    A = np.zeros((h, w))
    for off in (-7, -2, -1, 0, 3, 4, 5):
        for r in range(h):
            c = r + off
            if 0 <= c < w and rng.random() < 0.9:
                A[r, c] = rng.choice([-3, -2, -1, 1, 2, 3])
    rows, cols = np.nonzero(A)
    vals = A[rows, cols].astype(np.float32)
    coo = M.CooMatrix.from_arrays(rows.astype(np.int32), cols.astype(np.int32), vals, w, h)
    d = M.coo2dia(coo)
    b = M.coo2bdia(coo)
    assert list(b.offsets) == [-7, -2, 3] and list(b.bptr) == [0, 1, 4, 7]
    n = min(h, w)
    assert not np.array_equal(A[:n, :n], A[:n, :n].T)
    x = rng.integers(-2, 3, w).astype(np.float32)
    y0 = rng.integers(-5, 6, h).astype(np.float32)
    yd, yb = _dia_oracles(d.offsets, d.data, d.s.dia_length_aligned, b.bptr, b.offsets, x, y0, w, h)
    ref = y0 + A @ x
    assert np.array_equal(yd, ref) and np.array_equal(yb, ref)


def test_coo_spmv_oracle():
    """oracle_coo_spmv (y += A x in COO order, A/include/spmv_serial.h:184-196):
    from y = 0 bitwise equal to oracle_spmv_only (itself bit-exact against the
    reference's compiled spmv_only); from a nonzero y exact on a
    non-square integer-valued matrix (a transposed operand or a slipped
    index fails), and within FP32 of the dense FP64 product on real values."""
    o = O.lib()
    P = lambda a, t=C.c_int: np.ascontiguousarray(a).ctypes.data_as(C.POINTER(t))  # noqa: E731
    F = C.c_float
    coo = M.gen_zipf_rows(3000, 1.8, 8.0, 500, 9)
    x = M.gen_vector(coo.width, 4)
    y = np.zeros(coo.height, np.float32)
    o.oracle_coo_spmv(coo.nnz, P(coo.row_id), P(coo.col_id), P(coo.data, F), P(x, F), P(y, F))
    ys = O.spmv_only(coo.row_id, coo.col_id, coo.data, x, coo.height)
    assert np.array_equal(y.view(np.uint32), ys.view(np.uint32))
    rng = np.random.default_rng(8)
    h, w = 37, 53
    A = np.where(rng.random((h, w)) < 0.2, rng.integers(-4, 5, (h, w)), 0).astype(np.float32)
    r, c = np.nonzero(A)
    xi = rng.integers(-3, 4, w).astype(np.float32)
    y0 = rng.integers(-9, 10, h).astype(np.float32)
    y = y0.copy()
    o.oracle_coo_spmv(len(r), P(r.astype(np.int32)), P(c.astype(np.int32)), P(A[r, c], F), P(xi, F), P(y, F))
    assert np.array_equal(y, y0 + A.astype(np.float64) @ xi)
    xr = rng.uniform(-1, 1, w).astype(np.float32)
    Ar = np.where(A != 0, rng.standard_normal((h, w)), 0).astype(np.float32)
    y = y0.copy()
    o.oracle_coo_spmv(len(r), P(r.astype(np.int32)), P(c.astype(np.int32)), P(Ar[r, c], F), P(xr, F), P(y, F))
    assert np.allclose(y, y0 + Ar.astype(np.float64) @ xr, rtol=0, atol=1e-5)


def test_project_omp_csr_matches_serial_and_f64():
    # the project CPU baseline (OpenMP over rows): each row summed in order,
    # so bitwise equal to the serial CSR oracle; within FP32 of the FP64 product
    coo = M.gen_zipf_rows(20000, 1.8, 16.0, 3000, 3)
    csr = M.coo2csr(coo)
    x = M.gen_vector(coo.width, 1)
    y, threads = O.csr_spmv_omp(csr.row_ptr, csr.col_id, csr.data, x)
    assert threads >= 1
    ys = np.zeros(coo.height, np.float32)
    O.lib().oracle_csr_spmv(coo.height, O._p(np.ascontiguousarray(csr.row_ptr), C.c_int),
                            O._p(np.ascontiguousarray(csr.col_id), C.c_int),
                            O._p(np.ascontiguousarray(csr.data), C.c_float), O._p(x, C.c_float), O._p(ys, C.c_float))
    assert np.array_equal(y, ys)
    y64 = O.spmv_f64(coo.row_id, coo.col_id, coo.data, x, coo.height)
    S = O.abs_row_sum(coo.row_id, coo.col_id, coo.data, x, coo.height)
    assert np.max(np.abs(y - y64) / np.maximum(S, 1e-30)) < 1e-5


def test_two_vec_compare_is_sign_blind():
    a = np.array([1.0, -2.0, 3.0], np.float32)
    b = np.array([1.0, 2.0, 3.5], np.float32)
    r = O.two_vec_compare(a, b)
    assert r["max_diff"] == pytest.approx(0.5) and r["distance"] == pytest.approx(np.sqrt(16.25))


def test_parity_report_catches_a_planted_error():
    coo = M.gen_zipf_rows(2000, 1.8, 8.0, 500, 2)
    x = M.gen_vector(coo.width, 1)
    y = O.spmv_only(coo.row_id, coo.col_id, coo.data, x, coo.height)
    assert O.parity_report(y, coo, x)["ok"]
    bad = y.copy()
    bad[17] += 1e-3 * max(1.0, abs(bad[17]))
    assert not O.parity_report(bad, coo, x)["ok"]


# ------------------------------------------------------------- ABI surface
def _header_symbols():
    text = ""
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            text += open(os.path.join(ROOT, "include", f)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(clspmv_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    syms = _header_symbols()
    assert len(syms) > 60
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    # and the Python binding declares nothing the header lacks
    assert set(L.exported_symbols()) <= set(syms)


def test_no_device_is_an_error_not_a_fallback():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    h = C.c_void_p()
    assert L.lib().clspmv_init(0, C.byref(h)) == L.E_NODEVICE


def test_set_option_ranges():
    # every documented option takes its documented values and refuses the
    # rest (no device needed); the defaults are restored afterwards
    setopt = L.lib().clspmv_set_option
    cases = {b"col_slice_bytes": ([256, 1 << 21], [0, 255], 1 << 21),
             b"xtile_blocks": ([1, 256, 1 << 20], [0, (1 << 20) + 1], 256),
             b"xtile_max_groups": ([1, 8], [0, 9], 8),
             b"bsr_lds_window": ([0, 1, 2], [-1, 3], 1),
             b"cocktail_value_codec": ([0, L.VAL_BF16X2, L.VAL_FP16_ROW, L.VAL_FP8_SLICE, L.VAL_I16_ROW],
                                       [-1, L.VAL_I16_ROW + 1], 0)}
    try:
        for name, (good, bad, _) in cases.items():
            for v in good:
                assert setopt(name, v) == L.OK, (name, v)
            for v in bad:
                assert setopt(name, v) == L.E_INVALID, (name, v)
        assert setopt(b"no_such_option", 1) == L.E_INVALID
        assert setopt(None, 1) == L.E_INVALID
    finally:
        for name, (_, _, default) in cases.items():
            assert setopt(name, default) == L.OK


# ------------------------------------------------------------- failure paths
def test_dia_refuses_too_much_padding():
    coo = M.CooMatrix.from_arrays(np.arange(100, dtype=np.int32), (np.arange(100) * 37 % 100).astype(np.int32),
                                  np.ones(100, np.float32), 100, 100)
    with pytest.raises(L.ClspmvError) as e:
        M.coo2dia(coo)
    assert e.value.code == L.E_PADDING


def test_block_formats_refuse_above_max_mem_obj():
    # 1,048,577 8x8 blocks exceed MAX_MEM_OBJ / (8*8*4) = 1,048,576
    nb = 1048577
    rows = (np.arange(nb, dtype=np.int64) * 8).astype(np.int32)
    coo = M.CooMatrix.from_arrays(rows, rows, np.ones(nb, np.float32), nb * 8, nb * 8)
    for conv in (lambda c: M.coo2b4csr(c, 8, 8), lambda c: M.coo2b4ell(c, 8, 8),
                 lambda c: M.coo2sbell(c, 8, 8, 32)):
        with pytest.raises(L.ClspmvError) as e:
            conv(coo)
        assert e.value.code == L.E_TOO_LARGE


def test_invalid_indices_rejected():
    coo = M.CooMatrix.from_arrays(np.array([0, 5]), np.array([0, 1]), np.ones(2, np.float32), 4, 4)
    with pytest.raises(L.ClspmvError) as e:
        M.coo2csr(coo)
    assert e.value.code == L.E_INVALID
    coo = M.CooMatrix.from_arrays(np.array([0, 1]), np.array([0, 1]), np.ones(2, np.float32), 4, 4)
    with pytest.raises(L.ClspmvError):
        M.coo2ell(coo, ellnum=0) and M.coo2b4csr(coo, 3, 2)  # bwidth must be 4 or 8


def test_empty_matrices():
    coo = M.CooMatrix.from_arrays(np.zeros(0), np.zeros(0), np.zeros(0, np.float32), 10, 10)
    assert M.coo2csr(coo).row_ptr.tolist() == [0] * 11
    assert M.coo2ell(coo).s.ell_num == 0
    assert M.coo2sell(coo, 32).slice_ptr.tolist() == [0, 0]
    assert M.coo2bdia(coo).s.bdia_band_num == 0
