"""GPU parity of the Cocktail path: analyze_matrix -> init_mat_kernels ->
do_spmv, compared with the reference's CPU SpMV restated in oracle/ (same
bound as the single-format tests), on matrices whose partition has one to
three parts; plus the y = A^k x iteration of A/src/eval.cpp:19-72."""
import contextlib

import numpy as np
import pytest

from clspmv_amd import _lib as L
from clspmv_amd import cocktail as CK
from clspmv_amd import matrix as M
from oracle import oracle as O
from tests import analyzer_cases as AC

pytestmark = pytest.mark.gpu

CASES = ["lap2d_256", "st27_24", "st27_40", "uni_20k_32", "zipf_20k", "blk3", "blk4", "cocktail_64k",
         "band15_rand8", "band7_rand2", "blk4_rand", "rand_ragged"]
GEN = dict(AC.CASES)


@pytest.fixture(scope="module")
def tables():
    return CK.BenchTables(CK.REFERENCE_TABLES)


@pytest.fixture(params=[1, 0], ids=["merged_flat", "separate_flat"])
def merge_flat(request):
    # part counts below assume one merged flat part (no split into row runs)
    L.check(L.lib().clspmv_set_option(b"cocktail_merge_flat", request.param), "set_option")
    L.check(L.lib().clspmv_set_option(b"cocktail_split_local", 0), "set_option")
    yield request.param
    L.check(L.lib().clspmv_set_option(b"cocktail_merge_flat", 1), "set_option")
    L.check(L.lib().clspmv_set_option(b"cocktail_split_local", 1), "set_option")


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("knobs", AC.KNOBS, ids=lambda k: f"m{k[0]}p{k[1]}f{k[2]}")
def test_cocktail_spmv(ctx, tables, merge_flat, case, knobs):
    coo = AC.make(GEN[case])
    mat = CK.analyze_matrix(coo.copy(), tables, knobs[0], bool(knobs[1]), bool(knobs[2]))
    x = M.gen_vector(coo.width, 21)
    y0 = M.gen_vector(coo.height, 22)
    gpu = CK.GpuCocktail(ctx, mat, x, y0)
    flat = [n for n in mat.parts() if n in ("sell", "ell", "csr", "coo")]
    expect = len(mat.parts()) - (len(flat) - 1 if merge_flat and len(flat) >= 2 else 0)
    assert len(gpu.part_infos()) == expect, (mat.parts(), gpu.part_infos())
    gpu.do_spmv(1)
    y = gpu.result_to_cpu()
    rep = O.parity_report(y, coo, x, y0)
    assert rep["ok"], (case, mat.parts(), rep)
    gpu.free()


@pytest.mark.parametrize("case", ["cocktail_64k", "band15_rand8", "zipf_20k"])
def test_cocktail_mi355x_tables(ctx, merge_flat, case):
    coo = AC.make(GEN[case])
    mat = CK.analyze_matrix(coo.copy(), CK.BenchTables(CK.MI355X_TABLES))
    x = M.gen_vector(coo.width, 23)
    y0 = M.gen_vector(coo.height, 24)
    gpu = CK.GpuCocktail(ctx, mat, x, y0)
    gpu.do_spmv(1)
    rep = O.parity_report(gpu.result_to_cpu(), coo, x, y0)
    assert rep["ok"], (case, mat.parts(), rep)
    gpu.free()


@contextlib.contextmanager
def _option(name, value, default):
    """clspmv_set_option(name, value) for the block, then back to default."""
    L.check(L.lib().clspmv_set_option(name, value), "set_option")
    try:
        yield
    finally:
        L.check(L.lib().clspmv_set_option(name, default), "set_option")


def _local_scattered(width, seed):
    """3 x 32768 rows, 8 entries each: the first and last 32768 rows within
    2000 columns of the diagonal, the middle ones uniform over the width."""
    rng = np.random.default_rng(seed)
    B = 32768
    rows = np.repeat(np.arange(3 * B, dtype=np.int64), 8)
    near = np.clip(rows + rng.integers(-2000, 2001, rows.size), 0, width - 1)
    far = rng.integers(0, width, rows.size)
    cols = np.where((rows >= B) & (rows < 2 * B), far, near)
    key = np.unique(rows * width + cols)
    r, c = (key // width).astype(np.int32), (key % width).astype(np.int32)
    v = rng.standard_normal(r.size).astype(np.float32)
    return M.CooMatrix.from_arrays(r, c, v, width, 3 * B)


@pytest.mark.parametrize("width", [1 << 20, 1 << 22], ids=["x_in_l2", "wide_x"])
def test_cocktail_split_local(ctx, tables, width):
    # wide x: the merged flat rows become three row-run parts (local,
    # scattered, local) launched with y offset to their first rows; same
    # product as one part.  x within L2: the rows, all of one length, are one
    # SELL-class run wherever their columns lie, so the partition's own flat
    # part stays
    coo = _local_scattered(width, 31)
    mat = CK.analyze_matrix(coo.copy(), tables)
    flat = [n for n in mat.parts() if n in ("sell", "ell", "csr", "coo")]
    assert flat
    other = len(mat.parts()) - len(flat)
    x = M.gen_vector(coo.width, 25)
    y0 = M.gen_vector(coo.height, 26)
    ys = {}
    for split in (1, 0):
        # (runs at every size: 786432 entries are under the one-part size)
        with _option(b"cocktail_split_local", split, 1), _option(b"cocktail_small_plan", 0, 1 << 22):
            gpu = CK.GpuCocktail(ctx, mat, x, y0)
        parts = gpu.part_infos()
        runs = 3 if split and width > 1 << 20 else 1
        assert len(parts) == other + runs, (split, mat.parts(), parts)
        if runs == 3:
            assert sorted(p["rows"] for p in parts[other:]) == [32768, 32768, 32768]
        gpu.do_spmv(1)
        ys[split] = gpu.result_to_cpu()
        rep = O.parity_report(ys[split], coo, x, y0)
        assert rep["ok"], (split, width, rep)
        gpu.free()


def _blocks_scattered(seed, B=32768):
    """2 x B rows of width 2^22: dense aligned 4 x 4 blocks (6 per block
    row, near the diagonal), then 10 uniform columns per row."""
    rng = np.random.default_rng(seed)
    W = 1 << 22
    br = np.arange(B // 4)
    bc = np.clip(br[:, None] + rng.integers(-1000, 1001, (B // 4, 6)), 0, W // 4 - 1)
    rows = [(br[:, None] * 4 + i).repeat(6, 1).ravel() for i in range(4) for _ in range(4)]
    cols = [(bc * 4 + j).ravel() for _ in range(4) for j in range(4)]
    r = np.concatenate(rows + [np.repeat(np.arange(B, 2 * B), 10)]).astype(np.int64)
    c = np.concatenate(cols + [rng.integers(0, W, 10 * B)]).astype(np.int64)
    key = np.unique(r * W + c)
    r, c = (key // W).astype(np.int32), (key % W).astype(np.int32)
    return M.CooMatrix.from_arrays(r, c, rng.standard_normal(r.size).astype(np.float32), W, 2 * B)


def test_cocktail_runs_by_shape(ctx, tables):
    # local rows of dense 4 x 4 blocks run as BSR (a run of 786432 entries,
    # under 2^21: 2 x 2 blocks with lanes per block row), local rows of
    # equal length as SELL, scattered rows x-tiled or column-sliced
    cases = [(_blocks_scattered(41), ["bsr22_L", "scattered"]),
             (_local_scattered(1 << 22, 43), ["sell", "scattered", "sell"])]
    for coo, kinds in cases:
        mat = CK.analyze_matrix(coo.copy(), tables)
        assert mat.parts() == ["ell"], mat.parts()
        x = M.gen_vector(coo.width, 27)
        y0 = M.gen_vector(coo.height, 28)
        with _option(b"cocktail_small_plan", 0, 1 << 22):  # runs at every size
            gpu = CK.GpuCocktail(ctx, mat, x, y0)
        names = [p["kernel"] for p in gpu.part_infos()]
        assert len(names) == len(kinds), names
        for name, kind in zip(names, kinds):
            if kind == "scattered":
                assert name.startswith(("xtile", "csr_sliced")), names
            else:
                assert name.startswith(kind), names
        gpu.do_spmv(1)
        rep = O.parity_report(gpu.result_to_cpu(), coo, x, y0)
        assert rep["ok"], (names, rep)
        gpu.free()


def test_power_iteration(ctx, tables):
    # y = A^k x pattern of the reference's example driver: each do_spmv
    # accumulates into y, then y is copied into x on the device, so
    # y_{k+1} = (A + I) y_k.  Checked against an FP64 iteration.
    coo = M.gen_laplacian2d(48, 48)
    vals = coo.data * np.float32(0.25)
    coo = M.CooMatrix.from_arrays(coo.row_id, coo.col_id, vals, coo.width, coo.height)
    mat = CK.analyze_matrix(coo.copy(), tables)
    x = M.gen_vector(coo.width, 5)
    gpu = CK.GpuCocktail(ctx, mat, x, np.zeros(coo.height, np.float32))
    k = 10
    y = CK.power_iterate(ctx, gpu, k)
    v = x.astype(np.float64)
    yc = np.zeros(coo.height)
    for _ in range(k):
        yc = yc + O.spmv_f64(coo.row_id, coo.col_id, coo.data, v.astype(np.float32), coo.height)
        v = yc.copy()
        # the device iterates in FP32: round the vector as it does
        v = v.astype(np.float32).astype(np.float64)
        yc = v.copy()
    scale = np.maximum(np.abs(yc), 1e-3 * np.abs(yc).max())
    assert np.max(np.abs(y - yc) / scale) < 1e-4
    gpu.free()


def test_power_iteration_device_equals_host_loop(ctx, tables):
    # the device-resident iteration (x := y by a kernel; direct launches or
    # captured graphs) takes bitwise the same steps as the reference's
    # host-driven loop (do_spmv, then cpy_vector_from_gpu), also across graph
    # sizes (70 iterations = a 64-iteration graph and a 6-iteration graph)
    coo = AC.make(GEN["cocktail_64k"])
    vals = (coo.data * np.float32(0.05)).astype(np.float32)
    coo = M.CooMatrix.from_arrays(coo.row_id, coo.col_id, vals, coo.width, coo.height)
    mat = CK.analyze_matrix(coo.copy(), tables)
    x = M.gen_vector(coo.width, 31)
    y0 = np.zeros(coo.height, np.float32)
    out = {}
    for graph in (0, 1):
        L.check(L.lib().clspmv_set_option(b"cocktail_graph", graph), "set_option")
        try:
            for k in (3, 70):
                a = CK.GpuCocktail(ctx, mat, x, y0)
                b = CK.GpuCocktail(ctx, mat, x, y0)
                out["dev", k] = CK.power_iterate(ctx, a, k)
                out["host", k] = CK.power_iterate_host(ctx, b, k)
                assert np.array_equal(out["dev", k].view(np.uint32), out["host", k].view(np.uint32)), (graph, k)
                a.free()
                b.free()
        finally:
            L.check(L.lib().clspmv_set_option(b"cocktail_graph", 0), "set_option")
    # the first step is y = 0 + A x: the oracle's product
    g = CK.GpuCocktail(ctx, mat, x, y0)
    rep = O.parity_report(CK.power_iterate(ctx, g, 1), coo, x, y0)
    assert rep["ok"], rep
    g.free()


def test_graph_replays_equal_direct_launches(ctx, tables):
    # do_spmv replayed from captured graphs equals direct launches bitwise,
    # replay after replay, with eager copies and host synchronisations in
    # between (each replay checked)
    coo = AC.make(GEN["band15_rand8"])
    mat = CK.analyze_matrix(coo.copy(), tables)
    x = M.gen_vector(coo.width, 33)
    y0 = M.gen_vector(coo.height, 34)
    g = CK.GpuCocktail(ctx, mat, x, y0)
    assert len(g.part_infos()) >= 2
    ref = []
    for n in (1, 2, 65):
        g.result_from_cpu(y0)
        g.do_spmv(n)
        ref.append(g.result_to_cpu())
    L.check(L.lib().clspmv_set_option(b"cocktail_graph", 1), "set_option")
    try:
        for rep in range(3):
            for n, r in zip((1, 2, 65), ref):
                g.result_from_cpu(y0)
                g.do_spmv(n)
                ctx.sync()
                assert np.array_equal(g.result_to_cpu().view(np.uint32), r.view(np.uint32)), (rep, n)
    finally:
        L.check(L.lib().clspmv_set_option(b"cocktail_graph", 0), "set_option")
    rep1 = O.parity_report(ref[0], coo, x, y0)
    assert rep1["ok"], rep1
    g.free()


def test_evaluate_report(ctx, tables):
    # evaluate: one checked do_spmv against the caller's CPU result, then a
    # timed loop; the reference's two_vec_compare fields are reported
    coo = AC.make(GEN["band15_rand8"])
    mat = CK.analyze_matrix(coo.copy(), tables)
    x = M.gen_vector(coo.width, 35)
    res = M.gen_vector(coo.height, 36)
    coores = (O.spmv_only(coo.row_id, coo.col_id, coo.data, x, coo.height) + res).astype(np.float32)
    r = CK.evaluate(ctx, mat, x, res, coores, ntimes=20)
    assert r["ntimes"] == 20 and r["nparts"] >= 1 and r["ms_per_spmv"] > 0 and r["gflops"] > 0
    # the reference's report is relative to |y| (ill-posed where y ~ 0), so
    # the absolute figures are the ones bounded here: FP32 rounding of O(1)
    # row sums
    assert r["max_diff"] < 1e-4 and r["distance"] < 1e-3, r
    assert r["max_diff_ratio"] <= r["max_rel_err"] + 1e-12
    # a wrong expected vector is reported, not hidden
    bad = coores.copy()
    bad[7] += 1.0
    r2 = CK.evaluate(ctx, mat, x, res, bad, ntimes=1)
    assert r2["max_diff_id"] == 7 and r2["max_diff"] > 0.5


def test_probes(ctx):
    # the copy probe checks what it copied; graph launches are timed
    import ctypes as C

    n = 1 << 22
    src = M.DeviceArray.from_numpy(ctx, M.gen_vector(n, 37))
    dst = M.DeviceArray(ctx, n)
    g = C.c_float()
    L.check(L.lib().clspmv_probe_copy(ctx.h, src.ptr, dst.ptr, 4 * n, 3, C.byref(g)), "probe_copy")
    assert g.value > 100
    assert np.array_equal(dst.download(), src.download())
    for nodes in (1, 16):
        us = C.c_float()
        L.check(L.lib().clspmv_probe_graph_launch(ctx.h, nodes, 20, C.byref(us)), "probe_graph_launch")
        assert 0 < us.value < 10000
    src.free()
    dst.free()


@pytest.mark.parametrize("case", ["cocktail_64k", "band15_rand8", "blk4_rand", "zipf_20k", "st27_24", "blk3",
                                  "lap2d_256"])
@pytest.mark.parametrize("tname", ["reference", "mi355x"])
def test_plan_kernels(ctx, case, tname):
    # the analyzer's MI355X plan uploaded as planned: one device part per
    # plan part, in the plan's formats, at oracle parity
    coo = AC.make(GEN[case])
    t = CK.BenchTables(CK.REFERENCE_TABLES if tname == "reference" else CK.MI355X_TABLES)
    plan = CK.Plan.analyze(coo.copy(), t)
    x = M.gen_vector(coo.width, 41)
    y0 = M.gen_vector(coo.height, 42)
    gpu = CK.GpuCocktail(ctx, plan, x, y0)
    infos = gpu.part_infos()
    parts = plan.parts()
    assert len(infos) == len(parts), (parts, infos)
    for q, inf in zip(parts, infos):
        assert inf["rows"] == q["row1"] - q["row0"] or q["source"] < 9, (q, inf)
        assert inf["format"] == q["format"] or q["method"] in (2, 4), (q, inf)
    gpu.do_spmv(1)
    rep = O.parity_report(gpu.result_to_cpu(), coo, x, y0)
    assert rep["ok"], (case, tname, rep)
    gpu.free()
    plan.free()


@pytest.mark.parametrize("rows", [16064, 73024, 262144])
def test_concurrent_runs_equal_one_stream(ctx, rows):
    # the row runs of a composite plan (SELL, sliced BSR, CSR) on up to four
    # streams give y bit for bit as on one stream, over several passes and
    # two calls, and at oracle parity after one pass
    coo = M.gen_cocktail_composite(rows, 1000 + rows)
    with _option(b"cocktail_small_plan", 0, 1 << 22):  # runs at every size
        plan = CK.Plan.analyze(coo.copy(), CK.BenchTables(CK.MI355X_TABLES))
    assert len(plan.parts()) >= 3 and all(q["source"] == 9 for q in plan.parts())
    x = M.gen_vector(coo.width, 43)
    y0 = M.gen_vector(coo.height, 44)
    out = {}
    try:
        for streams in (0, 1):
            L.check(L.lib().clspmv_set_option(b"cocktail_streams", streams), "set_option")
            gpu = CK.GpuCocktail(ctx, plan, x, y0)
            gpu.do_spmv(1)
            one = gpu.result_to_cpu()
            gpu.do_spmv(3)
            gpu.do_spmv(2)
            out[streams] = (one, gpu.result_to_cpu())
            gpu.free()
    finally:
        L.check(L.lib().clspmv_set_option(b"cocktail_streams", 1), "set_option")
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    rep = O.parity_report(out[1][0], coo, x, y0)
    assert rep["ok"], rep
    plan.free()
    coo.free()


def test_plan_parity_on_the_selector_family(ctx):
    # 60 matrices of the autotuner's family (stencils, Laplacians, uniform,
    # Zipf and block rows, composites; 2^10..2^16 rows) through the
    # analyzer's MI355X plan -- every run class and layout the plan picks
    # (DIA / BSR lanes, SELL, x-tiled, one part for small composites, runs
    # side by side on streams) -- at oracle parity after one pass
    from clspmv_amd import selector as S
    tables = CK.BenchTables(CK.MI355X_TABLES)
    rng = np.random.default_rng(11)
    seen = set()
    for i in range(60):
        kind, coo = S.random_matrix(i, rng, 10, 16)
        plan = CK.Plan.analyze(coo.copy(), tables)
        x = M.gen_vector(coo.width, 100 + i)
        y0 = M.gen_vector(coo.height, 200 + i)
        gpu = CK.GpuCocktail(ctx, plan, x, y0)
        names = [p["kernel"] for p in gpu.part_infos()]
        seen.update(n.split("_")[0] + ("_lanes" if "lanes" in n or "_L" in n else "") for n in names)
        gpu.do_spmv(1)
        rep = O.parity_report(gpu.result_to_cpu(), coo, x, y0)
        assert rep["ok"], (i, kind, coo.height, names, rep)
        gpu.free()
        plan.free()
        coo.free()
    # the family reaches the small-matrix layouts
    assert {"dia_lanes", "bsr22_lanes"} & seen, seen
