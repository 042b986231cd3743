"""GPU parity of the LDS x-tiled layout (method 4 of the ELL, SELL, CSR and
COO parts) against the CPU oracle, on shapes that exercise its edges: several
column slices with a ragged last one, row blocks split by entry count and by
the 32768-row cap, tiles with one and with many rounds of chunks (one row
block forced), rows longer than a chunk, empty rows and empty blocks, and
tiles whose entry count is an exact multiple of the chunk (no padding).
Results must also be bitwise repeatable."""
import ctypes as C

import numpy as np
import pytest

from clspmv_amd import _lib as L
from clspmv_amd import matrix as M
from oracle import oracle as O
from tests.helpers import gpu_spmv, rand_coo

pytestmark = pytest.mark.gpu
XT = 4


def _wide_long_rows():
    # 3 dense-ish long rows across all slices + sparse random rest
    rng = np.random.default_rng(9)
    h, w = 3000, 50001
    r = [np.full(20000, 7), np.full(9000, 1500), np.full(40000, 2999)]
    c = [np.sort(rng.choice(w, 20000, replace=False)), np.sort(rng.choice(w, 9000, replace=False)),
         np.sort(rng.choice(w, 40000, replace=False))]
    base = rand_coo(h, w, 30000, 10)
    rows = np.concatenate([base.row_id] + r).astype(np.int32)
    cols = np.concatenate([base.col_id] + c).astype(np.int32)
    key = rows.astype(np.int64) * w + cols
    _, keep = np.unique(key, return_index=True)
    rows, cols = rows[keep], cols[keep]
    o = np.lexsort((cols, rows))
    vals = rng.standard_normal(o.size).astype(np.float32)
    return M.CooMatrix.from_arrays(rows[o], cols[o], vals, w, h)


def _rand_wide(h, w, nnz, seed):
    """Random sorted COO over a wide column range (no dense key space)."""
    rng = np.random.default_rng(seed)
    key = np.unique(rng.integers(0, h, nnz).astype(np.int64) * w + rng.integers(0, w, nnz))
    r = (key // w).astype(np.int32)
    c = (key % w).astype(np.int32)
    return M.CooMatrix.from_arrays(r, c, rng.standard_normal(r.size).astype(np.float32), w, h)


def _run_lengths():
    """Rows of 1..97 entries (lengths (7 r) mod 97 + 1) in one column slice:
    segments of every length start at every lane of a chunk, crossing the
    16-lane rows of the wave (the chunk reduction's general path)."""
    rng = np.random.default_rng(16)
    rows, cols = [], []
    for r in range(700):
        n = (7 * r) % 97 + 1
        rows.append(np.full(n, r, np.int32))
        cols.append(np.sort(rng.choice(4000, n, replace=False)).astype(np.int32))
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    return M.CooMatrix.from_arrays(r, c, rng.standard_normal(r.size).astype(np.float32), 4000, 700)


CASES = [
    ("lap2d_32", lambda: M.gen_laplacian2d(32, 32)),
    ("run_lengths", _run_lengths),
    ("uniform_rows", lambda: M.gen_uniform_rows(3000, 17, 3)),
    ("zipf", lambda: M.gen_zipf_rows(5000, 1.8, 16.0, 3000, 11)),
    ("rand_ragged", lambda: rand_coo(333, 301, 4000, 1)),
    ("rand_sparse_empty_rows", lambda: rand_coo(517, 517, 90, 2)),
    ("one_row", lambda: rand_coo(1, 40, 25, 3)),
    ("wide_ragged_slices", lambda: rand_coo(5000, 100003, 200000, 4)),
    ("rows_over_block_cap", lambda: M.gen_uniform_rows(70000, 8, 5)),
    ("exact_chunks", lambda: M.gen_uniform_rows(4096, 32, 6)),
    ("wide_long_rows", _wide_long_rows),
    ("zipf_wide", lambda: M.gen_zipf_rows(60000, 1.8, 16.0, 20000, 12)),
    # wide x: many slices split into column groups, ragged last slice
    ("groups4", lambda: _rand_wide(20000, 1200000, 300000, 14)),
    ("groups8", lambda: _rand_wide(3000, 5000001, 100000, 15)),
]
IDS = [c[0] for c in CASES]


def _check(coo, y, x, y0, what):
    rep = O.parity_report(y, coo, x, y0)
    assert rep["ok"], f"{what}: {rep}"


def _convert(fmt, coo):
    if fmt == L.FMT_ELL:
        return M.coo2ell(coo)
    if fmt == L.FMT_SELL:
        return M.coo2sell(coo, 32)
    if fmt == L.FMT_CSR_VECTOR:
        return M.coo2csr(coo)
    return coo


@pytest.fixture
def xt_blocks():
    def set_blocks(n):
        L.check(L.lib().clspmv_set_option(b"xtile_blocks", n), "set_option")
    yield set_blocks
    set_blocks(256)


@pytest.fixture
def xt_groups():
    L.check(L.lib().clspmv_set_option(b"xtile_max_groups", 8), "set_option")
    yield
    L.check(L.lib().clspmv_set_option(b"xtile_max_groups", 8), "set_option")


def _groups(info):
    return int(info["kernel"].rsplit("_G", 1)[1])


@pytest.mark.parametrize("name,make", CASES, ids=IDS)
@pytest.mark.parametrize("fmt", [L.FMT_ELL, L.FMT_SELL, L.FMT_CSR_VECTOR, L.FMT_COO], ids=["ell", "sell", "csr", "coo"])
def test_xtile(ctx, xt_groups, name, make, fmt):
    coo = make()
    if fmt in (L.FMT_ELL, L.FMT_SELL) and name in ("zipf", "zipf_wide", "wide_long_rows"):
        pytest.skip("ELL/SELL padding of power-law rows is too large for a unit test")
    x = M.gen_vector(coo.width, 5)
    y0 = M.gen_vector(coo.height, 6)
    y, info = gpu_spmv(ctx, fmt, _convert(fmt, coo), x, y0, XT)
    assert info["method"] == XT and info["kernel"].startswith("xtile"), info
    if name.startswith("groups"):
        assert _groups(info) > 1, info
    _check(coo, y, x, y0, f"xtile fmt{fmt} {name}")


@pytest.mark.parametrize("nblocks", [1, 3, 1000])
@pytest.mark.parametrize("name", ["uniform_rows", "zipf", "wide_long_rows", "exact_chunks", "groups4", "groups8"])
@pytest.mark.parametrize("groups", [1, 8])
def test_xtile_block_counts(ctx, xt_blocks, nblocks, name, groups):
    # one row block gives tiles of many 16-chunk rounds
    coo = dict(CASES)[name]()
    xt_blocks(nblocks)
    L.check(L.lib().clspmv_set_option(b"xtile_max_groups", groups), "set_option")
    x = M.gen_vector(coo.width, 7)
    y0 = M.gen_vector(coo.height, 8)
    try:
        y, info = gpu_spmv(ctx, L.FMT_CSR_VECTOR, M.coo2csr(coo), x, y0, XT)
    finally:
        L.check(L.lib().clspmv_set_option(b"xtile_max_groups", 8), "set_option")
    assert _groups(info) <= groups
    _check(coo, y, x, y0, f"xtile blocks={nblocks} {name} {info['kernel']}")


@pytest.mark.parametrize("name", ["uniform_rows", "zipf", "exact_chunks", "groups4", "rows_over_block_cap"])
def test_xtile_scan_prio_same_bits(ctx, name):
    # the issue priority of the waves on scan chunks (xtile_scan_prio 0 / 1
    # / the layout's choice) changes no sum: y is the same bits
    coo = dict(CASES)[name]()
    x = M.gen_vector(coo.width, 23)
    y0 = M.gen_vector(coo.height, 24)
    csr = M.coo2csr(coo)
    ys = []
    try:
        for prio in (0, 1, -1):
            L.check(L.lib().clspmv_set_option(b"xtile_scan_prio", prio), "set_option")
            y, _ = gpu_spmv(ctx, L.FMT_CSR_VECTOR, csr, x, y0, XT)
            ys.append(y)
    finally:
        L.check(L.lib().clspmv_set_option(b"xtile_scan_prio", -1), "set_option")
    assert np.array_equal(ys[0].view(np.uint32), ys[1].view(np.uint32)), name
    assert np.array_equal(ys[0].view(np.uint32), ys[2].view(np.uint32)), name
    _check(coo, ys[1], x, y0, f"xtile scan prio 1 {name}")


def test_xtile_repeatable(ctx):
    coo = M.gen_zipf_rows(60000, 1.8, 16.0, 20000, 12)
    x = M.gen_vector(coo.width, 1)
    y0 = np.zeros(coo.height, np.float32)
    csr = M.coo2csr(coo)
    ya, _ = gpu_spmv(ctx, L.FMT_CSR_VECTOR, csr, x, y0, XT)
    yb, _ = gpu_spmv(ctx, L.FMT_CSR_VECTOR, csr, x, y0, XT)
    assert np.array_equal(ya.view(np.uint32), yb.view(np.uint32))


def test_xtile_accumulates(ctx):
    coo = rand_coo(2000, 40000, 30000, 13)
    x = M.gen_vector(coo.width, 2)
    y0 = M.gen_vector(coo.height, 3)
    y, _ = gpu_spmv(ctx, L.FMT_CSR_VECTOR, M.coo2csr(coo), x, y0, XT, ntimes=3)
    y1 = O.spmv_f64(coo.row_id, coo.col_id, coo.data, x, coo.height)
    want = y0.astype(np.float64) + 3 * y1
    S = O.abs_row_sum(coo.row_id, coo.col_id, coo.data, x, coo.height)
    assert np.max(np.abs(y - want) / np.maximum(np.abs(want) + 3 * S, 1e-30)) < 1e-5


def test_xtile_empty(ctx):
    coo = M.CooMatrix.from_arrays(np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0, np.float32), 100, 50)
    y0 = M.gen_vector(50, 1)
    y, _ = gpu_spmv(ctx, L.FMT_CSR_VECTOR, M.coo2csr(coo), M.gen_vector(100, 2), y0, XT)
    assert np.array_equal(y, y0)


def test_xtile_even_rows_when_balancing_costs_groups(ctx, xt_blocks):
    # 3 x 32768 Zipf rows with 16 target workgroups: balancing by entries
    # would need a fourth row block (and only 4 groups each); the rows are
    # cut evenly into 3 full blocks of 5 column groups instead
    coo = M.gen_zipf_rows(3 * 32768, 1.8, 16.0, 20000, 31)
    xt_blocks(16)
    x = M.gen_vector(coo.width, 9)
    y0 = M.gen_vector(coo.height, 10)
    y, info = gpu_spmv(ctx, L.FMT_CSR_VECTOR, M.coo2csr(coo), x, y0, XT)
    assert info["kernel"] == "xtile_B3_G5", info
    _check(coo, y, x, y0, f"xtile even rows {info['kernel']}")
