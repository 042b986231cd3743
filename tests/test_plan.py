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
