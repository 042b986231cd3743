"""GPU parity of every single-format kernel against the CPU oracle.

Each case uploads one format through the C ABI (clspmv_part_*), runs
y += A x on the MI355X and compares with the reference's CPU SpMV restated in
oracle/ (FP32, COO order) using the per-element bound of
oracle.parity_report: relative error <= 1e-5 against the reference-order
FP32 result (scale max(|y|, sum_j |a_ij x_j|)), or no worse than that result
against the FP64 oracle.  y starts non-zero to check the y += semantics.
"""
import numpy as np
import pytest

from clspmv_amd import _lib as L
from clspmv_amd import matrix as M
from oracle import oracle as O
from tests.helpers import gpu_spmv, rand_coo

pytestmark = pytest.mark.gpu


def _cases():
    cases = []
    cases.append(("lap2d_32", lambda: M.gen_laplacian2d(32, 32)))
    cases.append(("stencil27_12", lambda: M.gen_stencil27(12, 7)))
    cases.append(("uniform_rows", lambda: M.gen_uniform_rows(3000, 17, 3)))
    cases.append(("zipf", lambda: M.gen_zipf_rows(5000, 1.8, 16.0, 3000, 11)))
    cases.append(("blocks3", lambda: M.gen_block_rows(700, 3, 5, 40, 5)))
    cases.append(("rand_ragged", lambda: rand_coo(333, 301, 4000, 1)))
    cases.append(("rand_sparse_empty_rows", lambda: rand_coo(517, 517, 90, 2)))
    cases.append(("one_row", lambda: rand_coo(1, 40, 25, 3)))
    return cases


CASES = _cases()
IDS = [c[0] for c in CASES]


def _x_y(coo, seed=5):
    x = M.gen_vector(coo.width, seed)
    y0 = M.gen_vector(coo.height, seed + 1)
    return x, y0


def _check(coo, y, x, y0, what):
    rep = O.parity_report(y, coo, x, y0)
    assert rep["ok"], f"{what}: {rep}"


@pytest.mark.parametrize("name,make", CASES, ids=IDS)
@pytest.mark.parametrize("method", [0, 1])
def test_ell(ctx, name, make, method):
    coo = make()
    x, y0 = _x_y(coo)
    ell = M.coo2ell(coo)
    y, info = gpu_spmv(ctx, L.FMT_ELL, ell, x, y0, method)
    _check(coo, y, x, y0, f"ELL m{method} {name}")


@pytest.mark.parametrize("name,make", CASES, ids=IDS)
@pytest.mark.parametrize("S", [32, 64, 128])
def test_sell(ctx, name, make, S):
    coo = make()
    x, y0 = _x_y(coo)
    y, _ = gpu_spmv(ctx, L.FMT_SELL, M.coo2sell(coo, S), x, y0)
    _check(coo, y, x, y0, f"SELL S{S} {name}")


@pytest.mark.parametrize("name,make", CASES, ids=IDS)
def test_csr_scalar(ctx, name, make):
    coo = make()
    x, y0 = _x_y(coo)
    y, _ = gpu_spmv(ctx, L.FMT_CSR_SCALAR, M.coo2csr(coo), x, y0)
    _check(coo, y, x, y0, f"CSR scalar {name}")


@pytest.mark.parametrize("name,make", CASES, ids=IDS)
@pytest.mark.parametrize("method", [0, 1])
def test_csr_vector(ctx, name, make, method):
    coo = make()
    x, y0 = _x_y(coo)
    y, info = gpu_spmv(ctx, L.FMT_CSR_VECTOR, M.coo2csr(coo), x, y0, method)
    _check(coo, y, x, y0, f"CSR vector m{method} {name} {info['kernel']}")


def test_csr_adaptive_long_rows(ctx):
    # rows far longer than one 8192-nnz chunk exercise the FP64 carry path
    rng = np.random.default_rng(9)
    h, w = 64, 100000
    lens = np.array([0, 70000, 3, 8192, 8193, 20000] + [int(v) for v in rng.integers(0, 200, h - 6)])
    rows, cols = [], []
    for r, l in enumerate(lens):
        c = np.sort(rng.choice(w, size=int(l), replace=False))
        rows.append(np.full(int(l), r, np.int32))
        cols.append(c.astype(np.int32))
    rows = np.concatenate(rows)
    cols = np.concatenate(cols)
    vals = rng.standard_normal(rows.size).astype(np.float32)
    coo = M.CooMatrix.from_arrays(rows, cols, vals, w, h)
    x, y0 = _x_y(coo)
    for method in (0, 1):
        y, info = gpu_spmv(ctx, L.FMT_CSR_VECTOR, M.coo2csr(coo), x, y0, method)
        _check(coo, y, x, y0, f"CSR long rows m{method}")


@pytest.fixture
def small_slices():
    # force many column slices on small matrices (256 columns per slice)
    L.check(L.lib().clspmv_set_option(b"col_slice_bytes", 1024), "set_option")
    yield
    L.check(L.lib().clspmv_set_option(b"col_slice_bytes", 2 << 20), "set_option")


@pytest.mark.parametrize("name,make", CASES, ids=IDS)
def test_column_sliced(ctx, small_slices, name, make):
    coo = make()
    x, y0 = _x_y(coo)
    y, info = gpu_spmv(ctx, L.FMT_CSR_VECTOR, M.coo2csr(coo), x, y0, 2)
    _check(coo, y, x, y0, f"CSR sliced {name} {info['kernel']}")
    y, info = gpu_spmv(ctx, L.FMT_ELL, M.coo2ell(coo), x, y0, 2)
    _check(coo, y, x, y0, f"ELL sliced {name} {info['kernel']}")
    y, info = gpu_spmv(ctx, L.FMT_SELL, M.coo2sell(coo, 32), x, y0, 2)
    _check(coo, y, x, y0, f"SELL sliced {name} {info['kernel']}")
    y, info = gpu_spmv(ctx, L.FMT_COO, coo, x, y0, 2)
    assert info["kernel"].startswith("coo_sliced"), info
    _check(coo, y, x, y0, f"COO sliced {name} {info['kernel']}")


def test_column_sliced_long_rows(ctx):
    # rows long within several slices exercise the grouped FP64 carries
    rng = np.random.default_rng(19)
    h, w = 40, 200000
    lens = np.array([150000, 0, 30000, 70, 9000] + [int(v) for v in rng.integers(0, 300, h - 5)])
    rows = np.concatenate([np.full(int(l), r, np.int32) for r, l in enumerate(lens)])
    cols = np.concatenate([np.sort(rng.choice(w, size=int(l), replace=False)).astype(np.int32) for l in lens])
    vals = rng.standard_normal(rows.size).astype(np.float32)
    coo = M.CooMatrix.from_arrays(rows, cols, vals, w, h)
    x, y0 = _x_y(coo)
    L.check(L.lib().clspmv_set_option(b"col_slice_bytes", 200000), "set_option")
    try:
        y, info = gpu_spmv(ctx, L.FMT_CSR_VECTOR, M.coo2csr(coo), x, y0, 2)
    finally:
        L.check(L.lib().clspmv_set_option(b"col_slice_bytes", 2 << 20), "set_option")
    _check(coo, y, x, y0, f"CSR sliced long rows {info['kernel']}")


@pytest.mark.parametrize("name,make", CASES, ids=IDS)
def test_coo(ctx, name, make):
    coo = make()
    x, y0 = _x_y(coo)
    y, _ = gpu_spmv(ctx, L.FMT_COO, coo, x, y0)
    _check(coo, y, x, y0, f"COO {name}")


def test_coo_long_rows_and_boundaries(ctx):
    # rows whose segments end exactly on lane (8) and chunk (512) boundaries
    lens = [8, 8, 16, 512, 1, 7, 504, 1024, 3, 5000, 0, 0, 9, 511, 513]
    rows = np.concatenate([np.full(l, r, np.int32) for r, l in enumerate(lens)])
    cols = np.concatenate([np.arange(l, dtype=np.int32) for l in lens])
    rng = np.random.default_rng(4)
    vals = rng.standard_normal(rows.size).astype(np.float32)
    coo = M.CooMatrix.from_arrays(rows, cols, vals, 6000, len(lens))
    x, y0 = _x_y(coo)
    y, _ = gpu_spmv(ctx, L.FMT_COO, coo, x, y0)
    _check(coo, y, x, y0, "COO boundaries")


@pytest.mark.parametrize("name,make", [c for c in CASES if c[0] in ("lap2d_32", "stencil27_12", "one_row")],
                         ids=["lap2d_32", "stencil27_12", "one_row"])
@pytest.mark.parametrize("method", [0, 1])
def test_dia_bdia(ctx, name, make, method):
    coo = make()
    x, y0 = _x_y(coo)
    if coo.width == coo.height:
        y, _ = gpu_spmv(ctx, L.FMT_DIA, M.coo2dia(coo), x, y0, method)
        _check(coo, y, x, y0, f"DIA m{method} {name}")
        y, _ = gpu_spmv(ctx, L.FMT_BDIA, M.coo2bdia(coo), x, y0, method)
        _check(coo, y, x, y0, f"BDIA m{method} {name}")


@pytest.mark.parametrize("name,make", [("lap2d_128", lambda: M.gen_laplacian2d(128, 128)),
                                       ("lap2d_129x127", lambda: M.gen_laplacian2d(129, 127)),
                                       ("stencil27_24", lambda: M.gen_stencil27(24, 5))],
                         ids=["lap2d_128", "lap2d_129x127", "stencil27_24"])
@pytest.mark.parametrize("method", [0, 1])
def test_dia_interior_workgroups(ctx, name, make, method):
    # enough rows that the four-row kernel has interior workgroups (every
    # column in x) between boundary ones, and a row count not divisible by 4
    coo = make()
    x, y0 = _x_y(coo)
    y, info = gpu_spmv(ctx, L.FMT_DIA, M.coo2dia(coo), x, y0, method)
    assert info["kernel"] == ("dia_row4" if method else "dia_row")
    _check(coo, y, x, y0, f"DIA m{method} {name}")


SHAPES = [(bh, bw) for bw in (4, 8) for bh in (1, 2, 4, 8)]


@pytest.mark.parametrize("bh,bw", SHAPES)
@pytest.mark.parametrize("name", ["blocks3", "rand_ragged", "lap2d_32"])
def test_block_formats(ctx, bh, bw, name):
    coo = dict(CASES)[name]()
    x, y0 = _x_y(coo)
    y, _ = gpu_spmv(ctx, L.FMT_BELL, M.coo2b4ell(coo, bw, bh), x, y0)
    _check(coo, y, x, y0, f"BELL {bh}x{bw} {name}")
    y, _ = gpu_spmv(ctx, L.FMT_SBELL, M.coo2sbell(coo, bw, bh, 32), x, y0)
    _check(coo, y, x, y0, f"SBELL {bh}x{bw} {name}")
    y, _ = gpu_spmv(ctx, L.FMT_BCSR, M.coo2b4csr(coo, bw, bh), x, y0)
    _check(coo, y, x, y0, f"BCSR {bh}x{bw} {name}")


@pytest.mark.parametrize("b", [2, 3, 4, 5, 6])
@pytest.mark.parametrize("method", [0, 1])
def test_bsr_native(ctx, b, method):
    coo = M.gen_block_rows(500, b, 6, 30, 8)
    x, y0 = _x_y(coo)
    y, info = gpu_spmv(ctx, L.FMT_BSR, M.coo2bsr(coo, b, b), x, y0, method)
    _check(coo, y, x, y0, f"BSR {b}x{b} {info['kernel']}")


@pytest.mark.parametrize("method", [0, 1])
def test_bsr_ragged(ctx, method):
    # unaligned shape (last block row/column partial) and uneven block rows
    coo = rand_coo(301, 299, 5000, 12)
    x, y0 = _x_y(coo)
    y, info = gpu_spmv(ctx, L.FMT_BSR, M.coo2bsr(coo, 3, 3), x, y0, method)
    _check(coo, y, x, y0, f"BSR ragged {info['kernel']}")


def test_bsr_sliced_int16_steps_and_fallback(ctx):
    # FP32 sliced BSR stores block columns as int16 steps within a block row
    # when every step fits; a step of 100000 columns keeps int32 columns
    near = M.gen_block_rows(3000, 3, 6, 500, 22)
    far_r = np.array([0, 0, 0, 1, 1, 1], np.int32)
    far_c = np.array([0, 1, 100000, 0, 1, 100001], np.int32)
    far = M.CooMatrix.from_arrays(far_r, far_c, np.arange(1, 7, dtype=np.float32), 100004, 2)
    for coo, tag in ((near, "_d16"), (far, "")):
        x, y0 = _x_y(coo)
        y, info = gpu_spmv(ctx, L.FMT_BSR, M.coo2bsr(coo, 3, 3), x, y0, 0)
        assert info["kernel"].startswith("bsr33_sliced" + tag) and info["method"] == 0, info
        assert ("_d16" in info["kernel"]) == bool(tag), info
        _check(coo, y, x, y0, f"BSR int16 steps {info['kernel']}")


def _blocks_coo(blocks, brows, cols, b=2, seed=0):
    """COO of dense b x b blocks {block row: [first column, ...]}."""
    rng = np.random.default_rng(seed)
    r, c = [], []
    for br, firsts in blocks.items():
        for c0 in firsts:
            for i in range(b):
                for j in range(b):
                    r.append(br * b + i)
                    c.append(c0 + j)
    r, c = np.array(r, np.int32), np.array(c, np.int32)
    o = np.lexsort((c, r))
    v = rng.standard_normal(r.size).astype(np.float32)
    return M.CooMatrix.from_arrays(r[o], c[o], v, cols, brows * b)


def _swap_blocks(bsr, a, b):
    """Swap blocks a and b of the BSR arrays (same matrix, other order)."""
    col, dat, st = bsr.col_id, bsr.data, bsr.s.bsr_block_stride
    col[a], col[b] = col[b], col[a]
    for q in range(bsr.s.bsr_bwidth * bsr.s.bsr_bheight):
        dat[q * st + a], dat[q * st + b] = dat[q * st + b], dat[q * st + a]


# (name, blocks per block row, block rows, swap the first two blocks of block
# row 0, int16 steps expected): steps of +-32768 are the int16 limits (2x2
# block columns are even, so +32766 is the largest positive step)
BSR_STEP_CASES = [
    ("step_32766", {0: [0, 32766], 1: [2, 4]}, 2, False, True),
    ("step_32768", {0: [0, 32768], 1: [2, 4]}, 2, False, False),
    ("step_minus_32768", {0: [7232, 40000], 1: [2, 4]}, 2, True, True),
    ("step_minus_32770", {0: [7230, 40000], 1: [2, 4]}, 2, True, False),
    ("empty_slice", {**{r: [2 * r, 2 * r + 30000] for r in range(64)},
                     **{r: [2 * r - 256, 2 * r + 32000] for r in range(128, 200)}}, 200, False, True),
]


@pytest.mark.parametrize("name,blocks,brows,swap,d16", BSR_STEP_CASES, ids=[c[0] for c in BSR_STEP_CASES])
def test_bsr_sliced_int16_step_limits(ctx, name, blocks, brows, swap, d16):
    # int16 column steps hold exactly the steps in [-32768, 32767]; a step
    # outside keeps int32 columns; a block row stored out of column order
    # (negative steps) and a slice of 64 empty block rows are handled
    coo = _blocks_coo(blocks, brows, 70000, seed=len(name))
    bsr = M.coo2bsr(coo, 2, 2)
    if swap:
        _swap_blocks(bsr, 0, 1)
        assert bsr.col_id[0] > bsr.col_id[1]
    x, y0 = _x_y(coo)
    y, info = gpu_spmv(ctx, L.FMT_BSR, bsr, x, y0, 0)
    assert info["kernel"].startswith("bsr22_sliced") and ("_d16" in info["kernel"]) == d16, (name, info)
    _check(coo, y, x, y0, f"BSR int16 step limits {name} {info['kernel']}")


@pytest.mark.parametrize("lanes", [2, 4, 8, 16])
def test_bsr_explicit_lanes(ctx, lanes):
    # method L = lanes per block row (16 shares its number with no other kind)
    coo = M.gen_block_rows(400, 3, 9, 60, 21)
    x, y0 = _x_y(coo)
    y, info = gpu_spmv(ctx, L.FMT_BSR, M.coo2bsr(coo, 3, 3), x, y0, lanes)
    assert info["kernel"] == f"bsr33_L{lanes}"
    _check(coo, y, x, y0, f"BSR {info['kernel']}")


def _bsr_window(flag):
    L.check(L.lib().clspmv_set_option(b"bsr_lds_window", flag), "set_option")


def _drop_rows(coo, lo, hi):
    keep = (coo.row_id < lo) | (coo.row_id >= hi)
    return M.CooMatrix.from_arrays(coo.row_id[keep].copy(), coo.col_id[keep].copy(), coo.data[keep].copy(),
                                   coo.width, coo.height)


# wide bands (x windows of 10k-30k columns per workgroup of 1024 block rows)
BSR_WIN_CASES = [("band_5_workgroups", lambda: M.gen_block_rows(5000, 3, 9, 2000, 31)),
                 ("ragged_uniform", lambda: rand_coo(3001, 20000, 40000, 14)),
                 ("empty_block_rows", lambda: _drop_rows(M.gen_block_rows(6000, 3, 7, 2000, 32), 3000, 4700)),
                 ("one_block_row", lambda: rand_coo(3, 30000, 40, 15))]


@pytest.mark.parametrize("b", [2, 3, 4])
@pytest.mark.parametrize("name,make", BSR_WIN_CASES, ids=[c[0] for c in BSR_WIN_CASES])
def test_bsr_lds_window_matches_plain(ctx, b, name, make):
    # the LDS-windowed sliced kernel does each lane's operations in the same
    # order as the plain one: bitwise equal results, both at oracle parity
    # (option 2: windows wherever they fit, these matrices being smaller
    # than the 128 windows the default rule asks for)
    coo = make()
    bsr = M.coo2bsr(coo, b, b)
    x, y0 = _x_y(coo)
    try:
        _bsr_window(2)
        yw, iw = gpu_spmv(ctx, L.FMT_BSR, bsr, x, y0, 0)
        _bsr_window(0)
        yp, ip = gpu_spmv(ctx, L.FMT_BSR, bsr, x, y0, 0)
    finally:
        _bsr_window(1)
    # names: bsr<b><b>_sliced[_d16][_xw] (int16 column steps where they fit)
    assert iw["kernel"].startswith(f"bsr{b}{b}_sliced") and iw["kernel"].endswith("_xw"), iw
    assert ip["kernel"].startswith(f"bsr{b}{b}_sliced") and not ip["kernel"].endswith("_xw"), ip
    assert np.array_equal(yw.view(np.uint32), yp.view(np.uint32)), name
    _check(coo, yw, x, y0, f"BSR {name} {iw['kernel']}")


@pytest.mark.parametrize("name,make", [("wider_than_lds", lambda: rand_coo(2000, 100000, 30000, 16)),
                                       ("narrow_band", lambda: M.gen_block_rows(5000, 2, 9, 10, 17))])
def test_bsr_window_out_of_range_uses_plain_kernel(ctx, name, make):
    # columns spread over more x than a workgroup's LDS holds, or a band so
    # narrow that the plain kernel's gathers are already local
    coo = make()
    x, y0 = _x_y(coo)
    y, info = gpu_spmv(ctx, L.FMT_BSR, M.coo2bsr(coo, 2, 2), x, y0, 0)
    assert info["kernel"].startswith("bsr22_sliced") and not info["kernel"].endswith("_xw"), info
    _check(coo, y, x, y0, f"BSR {name} {info['kernel']}")


def test_repeat_accumulates(ctx):
    # ntimes launches accumulate: y = y0 + 3 A x
    coo = M.gen_laplacian2d(16, 16)
    x = np.ones(coo.width, np.float32)
    y0 = np.zeros(coo.height, np.float32)
    y, _ = gpu_spmv(ctx, L.FMT_ELL, M.coo2ell(coo), x, y0, 0, ntimes=3)
    ref = O.spmv_only(coo.row_id, coo.col_id, coo.data, x, coo.height)
    assert np.array_equal(y, 3 * ref)
