"""The MI355X-mode plan (clspmv_analyze_plan) on the CPU: the part lists of
the analyzer cases and of the 4M-row Cocktail composite equal the committed
fixture (tests/golden/plan_mi355x.json); the composite's plan follows from
its generator's definition (band rows -> SELL, dense 4x4 block rows ->
sliced BSR, Zipf rows over the whole width -> the LDS x-tiled layout); and
every plan covers each nonzero once (row runs contiguous over all rows)."""
import json
import os

import pytest

from clspmv_amd import _lib as L
from clspmv_amd import cocktail as CK
from clspmv_amd import workloads as W
from tests import analyzer_cases as AC

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "plan_mi355x.json")))
FIELDS = GOLD["fields"]


@pytest.fixture(scope="module")
def tables():
    return {"reference": CK.BenchTables(CK.REFERENCE_TABLES), "mi355x": CK.BenchTables(CK.MI355X_TABLES)}


def _check_cover(parts, coo):
    runs = [q for q in parts if q["source"] == 9]
    if runs:
        assert runs[0]["row0"] == 0 and runs[-1]["row1"] == coo.height
        for a, b in zip(runs, runs[1:]):
            assert a["row1"] == b["row0"] and a["row0"] < a["row1"]
    assert sum(q["nnz"] for q in parts) == coo.nnz


@pytest.mark.parametrize("case", [c[0] for c in AC.CASES])
@pytest.mark.parametrize("tname", ["reference", "mi355x"])
def test_plan_matches_fixture(tables, case, tname):
    coo = AC.make(dict(AC.CASES)[case])
    p = CK.Plan.analyze(coo.copy(), tables[tname])
    parts = p.parts()
    assert [[q[f] for f in FIELDS] for q in parts] == GOLD["plans"][f"{case}/{tname}"]
    _check_cover(parts, coo)
    p.free()


@pytest.mark.parametrize("tname", ["reference", "mi355x"])
def test_composite_plan_follows_its_structure(tables, tname):
    coo = W.make_coo("c2")
    p = CK.Plan.analyze(coo.copy(), tables[tname])
    parts = p.parts()
    assert [[q[f] for f in FIELDS] for q in parts] == GOLD["plans"][f"c2/{tname}"]
    M1 = 1 << 20
    assert [(q["format"], q["row0"], q["row1"]) for q in parts] == [
        (L.FMT_SELL, 0, M1), (L.FMT_BSR, M1, 2 * M1), (L.FMT_CSR_VECTOR, 2 * M1, 4 * M1)]
    assert parts[2]["method"] == 4  # LDS x-tiled
    assert [q["block"] for q in parts] == [0, 4, 0]  # sliced BSR 4 x 4
    assert parts[1]["nnz"] == M1 // 4 * 6 * 16  # 6 dense 4 x 4 blocks per block row
    _check_cover(parts, coo)
    p.free()
    coo.free()


def test_plan_of_a_callers_partition_equals_analyze_plan(tables):
    coo = AC.make(dict(AC.CASES)["cocktail_64k"])
    mat = CK.analyze_matrix(coo.copy(), tables["mi355x"])
    a = CK.Plan.of(mat).parts()
    b = CK.Plan.analyze(coo.copy(), tables["mi355x"]).parts()
    assert a == b


def test_band_and_block_runs(tables):
    # with the MI355X tables the analyzer leaves stencils and 3 x 3 block
    # matrices in its flat parts; the plan runs their rows as DIA (27 or 5
    # diagonals, filled) and as whole-matrix sliced BSR 3 x 3
    for case, fmt in (("st27_24", L.FMT_DIA), ("st27_40", L.FMT_DIA), ("lap2d_64", L.FMT_DIA), ("blk3", L.FMT_BSR)):
        coo = AC.make(dict(AC.CASES)[case])
        parts = CK.Plan.analyze(coo.copy(), tables["mi355x"]).parts()
        assert [(q["source"], q["format"], q["row0"], q["row1"]) for q in parts] == [(9, fmt, 0, coo.height)], case
        assert parts[0]["block"] == (3 if fmt == L.FMT_BSR else 0), parts


@pytest.mark.parametrize("rows", [16064, 23824, 73024, 358448])
def test_composite_runs_follow_unaligned_boundaries(tables, rows):
    # band rows [0, q1), 4 x 4 block rows [q1, q2), Zipf rows [q2, rows) with
    # q1, q2 not on a run block: each boundary lands on the 128-row step
    # next to it, the step holding both kinds of rows staying out of the BSR
    # run, so that run holds block rows only (24 entries each)
    from clspmv_amd import matrix as M
    coo = M.gen_cocktail_composite(rows, 1000 + rows)
    L.check(L.lib().clspmv_set_option(b"cocktail_small_plan", 0), "set_option")  # runs at every size
    try:
        p = CK.Plan.analyze(coo.copy(), tables["mi355x"])
    finally:
        L.check(L.lib().clspmv_set_option(b"cocktail_small_plan", 1 << 22), "set_option")
    parts = p.parts()
    q1, q2 = rows // 4, rows // 2
    assert [q["format"] for q in parts] == [L.FMT_SELL, L.FMT_BSR, L.FMT_CSR_VECTOR], parts
    bsr = parts[1]
    assert bsr["row0"] == -(-q1 // 128) * 128 and bsr["row1"] == q2 // 128 * 128, (q1, q2, parts)
    assert bsr["nnz"] == 24 * (bsr["row1"] - bsr["row0"])
    # 4 x 4 blocks: sliced from 2^21 entries, below with lanes per block row as 2 x 2 blocks
    assert (bsr["method"], bsr["block"]) == ((0, 4) if bsr["nnz"] >= 1 << 21 else (1, 2)), bsr
    _check_cover(parts, coo)
    p.free()
    coo.free()


def test_small_matrix_runs_as_one_part(tables):
    # under 2^22 entries a composite's runs become one adaptive CSR part;
    # from there on the runs stay (cocktail_small_plan sets the size)
    from clspmv_amd import matrix as M
    small = M.gen_cocktail_composite(73024, 7)
    assert small.nnz < 1 << 22
    parts = CK.Plan.analyze(small.copy(), tables["mi355x"]).parts()
    assert [(q["source"], q["format"], q["method"], q["row0"], q["row1"]) for q in parts] == [
        (9, L.FMT_CSR_VECTOR, 1, 0, small.height)]
    L.check(L.lib().clspmv_set_option(b"cocktail_small_plan", small.nnz + 1), "set_option")
    try:
        assert len(CK.Plan.analyze(small.copy(), tables["mi355x"]).parts()) == 1
        L.check(L.lib().clspmv_set_option(b"cocktail_small_plan", small.nnz), "set_option")
        assert len(CK.Plan.analyze(small.copy(), tables["mi355x"]).parts()) == 3
    finally:
        L.check(L.lib().clspmv_set_option(b"cocktail_small_plan", 1 << 22), "set_option")
    small.free()
