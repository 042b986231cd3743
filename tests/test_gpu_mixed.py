"""Mixed-precision DIA and BSR parts against the FP64 oracle.  Error metric
per element: |y - y64| / max(|y64|, S) with S = sum_j |a_ij x_j| (plus |y0|):
bounded by 2e-5 for BF16 hi+lo (about 16 significant bits), 5e-4 for FP16
with a per-row scale (11 bits), 0.08 for FP8 E4M3 with a per-slice scale
(4 bits: half an ulp is 1/16) and, for int16 fixed point with a per-row
scale s = amax / 32767, by the codec's own bound sum_j (s/2) |x_j| / S plus
FP32 rounding.  int16 delta columns give results bitwise
equal to int32 columns, and fall back to int32 when a step does not fit."""
import numpy as np
import pytest

from clspmv_amd import _lib as L
from clspmv_amd import matrix as M
from oracle import oracle as O
from tests.helpers import rand_coo

pytestmark = pytest.mark.gpu
BOUND = {L.VAL_BF16X2: 2e-5, L.VAL_FP16_ROW: 5e-4, L.VAL_FP8_SLICE: 0.08, L.VAL_I16_ROW: 2e-5}
CODECS = [L.VAL_BF16X2, L.VAL_FP16_ROW, L.VAL_FP8_SLICE, L.VAL_I16_ROW]


def _run(ctx, part, coo, x, y0):
    dx = M.DeviceArray.from_numpy(ctx, x)
    dy = M.DeviceArray.from_numpy(ctx, y0)
    part.spmv(dx, dy, 1)
    y = dy.download()
    dx.free()
    dy.free()
    return y


def _err(y, coo, x, y0):
    y64 = O.spmv_f64(coo.row_id, coo.col_id, coo.data, x, coo.height) + y0.astype(np.float64)
    S = O.abs_row_sum(coo.row_id, coo.col_id, coo.data, x, coo.height) + np.abs(y0)
    return float(np.max(np.abs(y - y64) / np.maximum(np.maximum(np.abs(y64), S), 1e-30)))


def _row_amax(coo, group=1, window=0):
    """Largest |a_ij| of each row's scale group: the row (group 1), the
    aligned block row of `group` rows, or any row within `window` rows."""
    a = np.zeros(coo.height, np.float64)
    np.maximum.at(a, coo.row_id, np.abs(coo.data).astype(np.float64))
    if group > 1:
        n = (coo.height + group - 1) // group
        a = np.repeat(np.pad(a, (0, n * group - coo.height)).reshape(n, group).max(1), group)[:coo.height]
    for k in range(1, window + 1):
        a = np.maximum(a, np.maximum(np.pad(a[k:], (0, k)), np.pad(a[:-k], (k, 0))))
    return a


def _check(codec, y, coo, x, y0, group=1, window=0):
    """The codec's bound.  int16 rows: |y - y64| <= (s/2) sum_j |x_j| plus
    1e-5 (S + |y0|) for the FP32 arithmetic, s = amax / 32767 of the row's
    scale group; the others: the componentwise bound BOUND[codec]."""
    e = _err(y, coo, x, y0)
    if codec != L.VAL_I16_ROW:
        assert e <= BOUND[codec], (codec, e)
        return e
    y64 = O.spmv_f64(coo.row_id, coo.col_id, coo.data, x, coo.height) + y0.astype(np.float64)
    S = O.abs_row_sum(coo.row_id, coo.col_id, coo.data, x, coo.height) + np.abs(y0)
    xs = np.bincount(coo.row_id, np.abs(x[coo.col_id]).astype(np.float64), minlength=coo.height)
    bound = _row_amax(coo, group, window) / 32767 / 2 * xs + 1e-5 * S
    d = np.abs(y - y64)
    assert np.all(d <= bound), (int(np.argmax(d - bound)), float(np.max(d - bound)))
    return e


def _rect_band(h, w, offsets, seed):
    """h x w with full diagonals at the given column offsets (col = row + o)."""
    rng = np.random.default_rng(seed)
    r = np.concatenate([np.arange(h)[(np.arange(h) + o >= 0) & (np.arange(h) + o < w)] for o in offsets])
    c = np.concatenate([np.arange(h)[(np.arange(h) + o >= 0) & (np.arange(h) + o < w)] + o for o in offsets])
    o = np.lexsort((c, r))
    v = rng.standard_normal(r.size).astype(np.float32)
    return M.CooMatrix.from_arrays(r[o].astype(np.int32), c[o].astype(np.int32), v, w, h)


DIA_CASES = [("lap2d", lambda: M.gen_laplacian2d(64, 48)), ("stencil27", lambda: M.gen_stencil27(20, 3)),
             ("band_ragged", lambda: M.gen_stencil27(7, 4)), ("lap2d_ragged", lambda: M.gen_laplacian2d(45, 41)),
             ("rect_wide", lambda: _rect_band(700, 3000, [-3, 0, 5, 1000, 2299], 17))]


@pytest.mark.parametrize("codec", CODECS)
@pytest.mark.parametrize("name,make", DIA_CASES, ids=[c[0] for c in DIA_CASES])
def test_dia_mixed(ctx, codec, name, make):
    coo = make()
    x, y0 = M.gen_vector(coo.width, 1), M.gen_vector(coo.height, 2)
    part = M.MixedPart(ctx, L.FMT_DIA, M.coo2dia(coo, 256), codec)
    assert part.info["method"] == 16 + codec
    _check(codec, _run(ctx, part, coo, x, y0), coo, x, y0)
    part.free()


BSR_CASES = [("blocks2", lambda: M.gen_block_rows(900, 2, 7, 60, 5)), ("blocks3", lambda: M.gen_block_rows(700, 3, 9, 50, 6)),
             ("blocks4", lambda: M.gen_block_rows(500, 4, 5, 40, 7)), ("ragged", lambda: rand_coo(301, 299, 3000, 8))]


@pytest.mark.parametrize("codec", CODECS)
@pytest.mark.parametrize("b", [2, 3, 4])
@pytest.mark.parametrize("name,make", BSR_CASES, ids=[c[0] for c in BSR_CASES])
def test_bsr_mixed(ctx, codec, b, name, make):
    coo = make()
    x, y0 = M.gen_vector(coo.width, 3), M.gen_vector(coo.height, 4)
    bsr = M.coo2bsr(coo, b, b)
    p16 = M.MixedPart(ctx, L.FMT_BSR, bsr, codec, True)
    p32 = M.MixedPart(ctx, L.FMT_BSR, bsr, codec, False)
    assert p16.info["method"] == 16 + codec + 8 and p32.info["method"] == 16 + codec
    ya = _run(ctx, p16, coo, x, y0)
    yb = _run(ctx, p32, coo, x, y0)
    assert np.array_equal(ya.view(np.uint32), yb.view(np.uint32))
    _check(codec, ya, coo, x, y0, group=b)
    p16.free()
    p32.free()


@pytest.mark.parametrize("codec", CODECS)
def test_bsr_mixed_lds_window_matches_plain(ctx, codec):
    # five workgroups of 1024 block rows, x windows staged in LDS, against
    # the plain kernel (x read from global memory): bitwise equal
    coo = M.gen_block_rows(5000, 3, 9, 2000, 33)
    bsr = M.coo2bsr(coo, 3, 3)
    x, y0 = M.gen_vector(coo.width, 3), M.gen_vector(coo.height, 4)
    try:
        # windows wherever they fit (5 windows: fewer than the default rule's 128)
        L.check(L.lib().clspmv_set_option(b"bsr_lds_window", 2), "set_option")
        pw = M.MixedPart(ctx, L.FMT_BSR, bsr, codec, True)
        L.check(L.lib().clspmv_set_option(b"bsr_lds_window", 0), "set_option")
        pp = M.MixedPart(ctx, L.FMT_BSR, bsr, codec, True)
    finally:
        L.check(L.lib().clspmv_set_option(b"bsr_lds_window", 1), "set_option")
    assert pw.info["kernel"].endswith("_xw") and not pp.info["kernel"].endswith("_xw")
    yw, yp = _run(ctx, pw, coo, x, y0), _run(ctx, pp, coo, x, y0)
    assert np.array_equal(yw.view(np.uint32), yp.view(np.uint32))
    _check(codec, yw, coo, x, y0, group=3)
    pw.free()
    pp.free()


def test_bsr_delta_falls_back_on_wide_steps(ctx):
    # two blocks per block row 100000 columns apart: the step needs int32
    rows = np.array([0, 0, 1, 1], np.int32)
    cols = np.array([0, 100000, 1, 100001], np.int32)
    coo = M.CooMatrix.from_arrays(rows, cols, np.ones(4, np.float32), 100004, 2)
    part = M.MixedPart(ctx, L.FMT_BSR, M.coo2bsr(coo, 2, 2), L.VAL_FP16_ROW, True)
    assert part.info["method"] == 16 + L.VAL_FP16_ROW and "_d16" not in part.info["kernel"]
    x, y0 = M.gen_vector(coo.width, 5), np.zeros(2, np.float32)
    assert _err(_run(ctx, part, coo, x, y0), coo, x, y0) <= BOUND[L.VAL_FP16_ROW]
    part.free()


def _integer_valued(coo, seed):
    """The same pattern with values in {-8..8} \\ {0}: exact in BF16 and,
    after a power-of-two row scale, in FP16."""
    rng = np.random.default_rng(seed)
    v = rng.integers(1, 9, coo.nnz) * rng.choice([-1, 1], coo.nnz)
    return M.CooMatrix.from_arrays(coo.row_id.copy(), coo.col_id.copy(), v.astype(np.float32), coo.width, coo.height)


def _wide_long_rows():
    from tests.test_gpu_xtile import _wide_long_rows as make
    return make()


CSR_CASES = [("uniform_rows", lambda: M.gen_uniform_rows(3000, 17, 3)),
             ("zipf", lambda: M.gen_zipf_rows(5000, 1.8, 16.0, 3000, 11)),
             ("wide_ragged_slices", lambda: rand_coo(5000, 100003, 200000, 4)),
             ("rows_over_block_cap", lambda: M.gen_uniform_rows(40000, 8, 5)),
             ("empty_rows", lambda: rand_coo(517, 517, 90, 2)),
             ("wide_long_rows", _wide_long_rows)]


@pytest.mark.parametrize("codec", CODECS)
@pytest.mark.parametrize("name,make", CSR_CASES, ids=[c[0] for c in CSR_CASES])
def test_csr_mixed_xtile(ctx, codec, name, make):
    coo = make()
    x, y0 = M.gen_vector(coo.width, 6), M.gen_vector(coo.height, 7)
    part = M.MixedPart(ctx, L.FMT_CSR_VECTOR, M.coo2csr(coo), codec)
    assert part.info["method"] == 16 + codec and part.info["kernel"].startswith("xtile_")
    _check(codec, _run(ctx, part, coo, x, y0), coo, x, y0)
    part.free()


@pytest.mark.parametrize("codec", [L.VAL_BF16X2, L.VAL_FP16_ROW])
@pytest.mark.parametrize("name,make", CSR_CASES, ids=[c[0] for c in CSR_CASES])
def test_csr_mixed_exact_values_match_fp32(ctx, codec, name, make):
    # Values that both codecs hold exactly (FP16 after an exact power-of-two
    # scale): every product and partial sum is the FP32 path's times a power
    # of two, so y must be bitwise that of the FP32 x-tiled part.
    coo = _integer_valued(make(), 13)
    csr = M.coo2csr(coo)
    x, y0 = M.gen_vector(coo.width, 8), M.gen_vector(coo.height, 9)
    mixed = M.MixedPart(ctx, L.FMT_CSR_VECTOR, csr, codec)
    plain = M.Part(ctx, L.FMT_CSR_VECTOR, csr, 4)
    ya, yb = _run(ctx, mixed, coo, x, y0), _run(ctx, plain, coo, x, y0)
    assert np.array_equal(ya.view(np.uint32), yb.view(np.uint32)), (name, codec)
    mixed.free()
    plain.free()


@pytest.mark.parametrize("name,make", CSR_CASES, ids=[c[0] for c in CSR_CASES])
def test_csr_i16_exact_values_match_fp32(ctx, name, make):
    # Integer values in [-32767, 32767] with +-32767 in every nonempty row:
    # the row scale is exactly 1, the stored integers are the values, so y
    # must be bitwise that of the FP32 x-tiled part (pins the sign extension
    # of both 16-bit halves).
    coo = make()
    rng = np.random.default_rng(21)
    v = rng.integers(-32767, 32768, coo.nnz).astype(np.float32)
    first = np.flatnonzero(np.r_[True, coo.row_id[1:] != coo.row_id[:-1]]) if coo.nnz else np.zeros(0, np.int64)
    v[first] = np.where(rng.random(first.size) < 0.5, 32767.0, -32767.0)
    coo = M.CooMatrix.from_arrays(coo.row_id.copy(), coo.col_id.copy(), v, coo.width, coo.height)
    csr = M.coo2csr(coo)
    x, y0 = M.gen_vector(coo.width, 10), M.gen_vector(coo.height, 11)
    mixed = M.MixedPart(ctx, L.FMT_CSR_VECTOR, csr, L.VAL_I16_ROW)
    plain = M.Part(ctx, L.FMT_CSR_VECTOR, csr, 4)
    ya, yb = _run(ctx, mixed, coo, x, y0), _run(ctx, plain, coo, x, y0)
    assert np.array_equal(ya.view(np.uint32), yb.view(np.uint32)), name
    mixed.free()
    plain.free()


@pytest.mark.parametrize("fmt", ["dia", "bsr", "csr"])
def test_i16_refuses_non_finite(ctx, fmt):
    coo = M.gen_laplacian2d(32, 32)
    v = coo.data.copy()
    v[100] = np.inf
    coo = M.CooMatrix.from_arrays(coo.row_id.copy(), coo.col_id.copy(), v, coo.width, coo.height)
    mat = {"dia": lambda: M.coo2dia(coo, 256), "bsr": lambda: M.coo2bsr(coo, 2, 2), "csr": lambda: M.coo2csr(coo)}[fmt]()
    f = {"dia": L.FMT_DIA, "bsr": L.FMT_BSR, "csr": L.FMT_CSR_VECTOR}[fmt]
    with pytest.raises(RuntimeError, match="finite"):
        M.MixedPart(ctx, f, mat, L.VAL_I16_ROW)


def test_csr_mixed_stored_bytes(ctx):
    # 8 bytes per x-tiled entry in FP32 and BF16 hi+lo, 6 with FP16 or int16
    # (+ row scales), 5 with FP8 (+ one scale per 128-entry chunk)
    coo = M.gen_uniform_rows(4096, 32, 6)  # 128 chunks of 128 entries, no padding
    csr = M.coo2csr(coo)
    plain = M.Part(ctx, L.FMT_CSR_VECTOR, csr, 4)
    parts = {c: M.MixedPart(ctx, L.FMT_CSR_VECTOR, csr, c) for c in CODECS}
    base = plain.info["stored_bytes"] - 8 * coo.nnz
    assert parts[L.VAL_BF16X2].info["stored_bytes"] == plain.info["stored_bytes"]
    assert parts[L.VAL_FP16_ROW].info["stored_bytes"] == base + 6 * coo.nnz + 4 * coo.height
    assert parts[L.VAL_I16_ROW].info["stored_bytes"] == base + 6 * coo.nnz + 4 * coo.height
    assert parts[L.VAL_FP8_SLICE].info["stored_bytes"] == base + 5 * coo.nnz + 4 * (coo.nnz // 128)
    for p in list(parts.values()) + [plain]:
        p.free()


def _e4m3(q):
    """E4M3 (OCP) value of code q from its definition."""
    s, e, m = q >> 7, (q >> 3) & 15, q & 7
    v = 2.0 ** (e - 7) * (1 + m / 8) if e else m * 2.0 ** -9
    return -v if s else v


def test_e4m3_decode_every_code(ctx):
    # A 256-row diagonal (one FP8 slice) whose values are every finite E4M3
    # code plus 448, the largest, so that the slice scale is exactly 1: with
    # x = 1, y is the decoded codes themselves, bitwise.
    assert _e4m3(0x7e) == 448.0 and _e4m3(0x01) == 2.0 ** -9 and _e4m3(0x08) == 2.0 ** -6
    codes = [q for q in range(256) if q not in (0x7f, 0xff, 0x7e, 0xfe)]
    n = 256
    vals = np.full(n, 448.0, np.float32)
    vals[1:1 + len(codes)] = [_e4m3(q) for q in codes]
    vals[-1] = -448.0
    idx = np.arange(n, dtype=np.int32)
    coo = M.CooMatrix.from_arrays(idx, idx, vals, n, n)
    part = M.MixedPart(ctx, L.FMT_DIA, M.coo2dia(coo, 256), L.VAL_FP8_SLICE)
    y = _run(ctx, part, coo, np.ones(n, np.float32), np.zeros(n, np.float32))
    assert np.array_equal(y, vals)
    part.free()


@pytest.mark.parametrize("codec", CODECS)
def test_cocktail_value_codec(ctx, codec):
    # the Cocktail row-run parts with a mixed-precision variant (sliced BSR,
    # x-tiled) take the codec; the others stay FP32.  131072 block rows: a
    # BSR run of 3.1M entries, sliced (a run under 2^21 entries takes the
    # lanes layout, which has no codec variant)
    import ctypes as C

    from clspmv_amd import cocktail as CK
    from tests.test_gpu_cocktail import _blocks_scattered

    coo = _blocks_scattered(41, 131072)
    mat = CK.analyze_matrix(coo.copy(), CK.BenchTables(CK.REFERENCE_TABLES))
    x, y0 = M.gen_vector(coo.width, 29), M.gen_vector(coo.height, 30)
    L.check(L.lib().clspmv_set_option(b"cocktail_value_codec", C.c_longlong(codec)), "set_option")
    L.check(L.lib().clspmv_set_option(b"cocktail_codec_tol", C.c_longlong(0)), "set_option")  # no guard
    try:
        gpu = CK.GpuCocktail(ctx, mat, x, y0)
    finally:
        L.check(L.lib().clspmv_set_option(b"cocktail_value_codec", C.c_longlong(0)), "set_option")
        L.check(L.lib().clspmv_set_option(b"cocktail_codec_tol", C.c_longlong(100000)), "set_option")
    names = [p["kernel"] for p in gpu.part_infos()]
    assert names[0].startswith("bsr44_mixed_") and all("mixed" in n or n.startswith(("xtile_", "csr_sliced"))
                                                       for n in names), names
    gpu.do_spmv(1)
    _check(codec, gpu.result_to_cpu(), coo, x, y0, window=3)
    gpu.free()



MIXED_CASES = ([(L.FMT_DIA, n, mk) for n, mk in DIA_CASES[:3]] + [(L.FMT_BSR, n, mk) for n, mk in BSR_CASES[1:3]]
               + [(L.FMT_CSR_VECTOR, n, mk) for n, mk in CSR_CASES[:3]])


@pytest.mark.parametrize("codec", CODECS)
@pytest.mark.parametrize("fmt,name,make", MIXED_CASES, ids=[f"{c[0]}_{c[1]}" for c in MIXED_CASES])
def test_codec_bound_bounds_the_error(ctx, codec, fmt, name, make):
    # the encoder's codec_bound (largest relative error of a stored value)
    # bounds the componentwise error of y = A x for any x, up to the FP32
    # arithmetic (1e-5 of S); BF16 hi+lo holds ~16 bits, FP16 ~11 bits
    coo = make()
    mat = M.coo2dia(coo, 256) if fmt == L.FMT_DIA else M.coo2bsr(coo, 3, 3) if fmt == L.FMT_BSR else M.coo2csr(coo)
    part = M.MixedPart(ctx, fmt, mat, codec)
    bound = part.info["codec_bound"]
    assert 0 <= bound <= 1
    if codec == L.VAL_BF16X2:
        assert bound <= 2.0 ** -15, bound
    if codec == L.VAL_FP16_ROW:
        assert bound <= 2.0 ** -10, bound
    y0 = np.zeros(coo.height, np.float32)
    for x in (M.gen_vector(coo.width, 1), np.ones(coo.width, np.float32)):
        e = _err(_run(ctx, part, coo, x, y0), coo, x, y0)
        assert e <= bound + 1e-5, (e, bound)
    part.free()


def _with_values(coo, vals):
    return M.CooMatrix.from_arrays(coo.row_id.copy(), coo.col_id.copy(), np.asarray(vals, np.float32), coo.width,
                                   coo.height)


@pytest.mark.parametrize("codec", CODECS)
def test_cocktail_codec_guard(ctx, codec):
    # The Cocktail driver keeps a run's codec only when its codec_bound is
    # within 1e-4 (the default guard).  N(0,1) values: only BF16 hi+lo keeps
    # 1e-4 for every x; values in {+-1, 2, 4}: every codec holds them within
    # 1e-4; the same with one planted value 1e-6 in a block row: that run
    # falls back to FP32 except under BF16.  (The scattered run has a codec
    # variant only in the x-tiled layout; in the column-sliced one it stays
    # FP32.)
    import ctypes as C

    from clspmv_amd import cocktail as CK
    from tests.test_gpu_cocktail import _blocks_scattered

    base = _blocks_scattered(41, 131072)  # a sliced BSR run (see test_cocktail_value_codec)
    rng = np.random.default_rng(5)
    ints = _with_values(base, rng.choice([-4.0, -2.0, -1.0, 1.0, 2.0, 4.0], base.nnz))
    planted_vals = ints.data.copy()
    k = int(np.searchsorted(base.row_id, 4000))
    planted_vals[k] *= 1e-6
    planted = _with_values(base, planted_vals)
    tables = CK.BenchTables(CK.REFERENCE_TABLES)
    bf16 = codec == L.VAL_BF16X2
    for coo, expect in ((base, [bf16, bf16]), (ints, [True, True]), (planted, [bf16, True])):
        mat = CK.analyze_matrix(coo.copy(), tables)
        x, y0 = M.gen_vector(coo.width, 29), M.gen_vector(coo.height, 30)
        L.check(L.lib().clspmv_set_option(b"cocktail_value_codec", C.c_longlong(codec)), "set_option")
        try:
            gpu = CK.GpuCocktail(ctx, mat, x, y0)
        finally:
            L.check(L.lib().clspmv_set_option(b"cocktail_value_codec", C.c_longlong(0)), "set_option")
        infos = gpu.part_infos()
        kept = ["mixed" in p["kernel"] or "_" + {1: "bf16x2", 2: "fp16row", 3: "fp8slice", 4: "i16row"}[codec]
                in p["kernel"] for p in infos]
        if infos[1]["kernel"].startswith("csr_sliced"):
            expect = [expect[0], False]
        assert kept == expect, ([p["kernel"] for p in infos], expect)
        for p, kp in zip(infos, kept):
            assert (p["codec_bound"] <= 1e-4) if kp else p["codec_bound"] == 0.0, p
        gpu.do_spmv(1)
        e = _err(gpu.result_to_cpu(), coo, x, y0)
        assert e <= 1e-4 + 1e-5, ([p["kernel"] for p in infos], e)
        gpu.free()
        mat.free()
