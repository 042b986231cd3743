"""Writes tests/golden/plan_mi355x.json: the MI355X-mode plans
(clspmv_analyze_plan) of the analyzer cases and of the 4M-row Cocktail
composite, with the reference's tables and with the MI355X tables.  The
plans are this library's own decisions (row-run classes, layouts by the
cost rule), so the fixture is a regression pin: a change to the plan
builder shows up as a changed part list.

    python tests/golden/make_golden_plan.py
"""
from __future__ import annotations

import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from clspmv_amd import cocktail as CK  # noqa: E402
from clspmv_amd import workloads as W  # noqa: E402
from tests import analyzer_cases as AC  # noqa: E402

FIELDS = ("source", "format", "method", "row0", "row1", "nnz")


def plans():
    tables = {"reference": CK.BenchTables(CK.REFERENCE_TABLES), "mi355x": CK.BenchTables(CK.MI355X_TABLES)}
    cases = [(name, lambda g=gen: AC.make(g)) for name, gen in AC.CASES] + [("c2", lambda: W.make_coo("c2"))]
    out = {}
    for name, make in cases:
        coo = make()
        for tname, t in tables.items():
            p = CK.Plan.analyze(coo.copy(), t)
            out[f"{name}/{tname}"] = [[q[f] for f in FIELDS] for q in p.parts()]
            p.free()
        coo.free()
    return out


if __name__ == "__main__":
    with open(os.path.join(HERE, "plan_mi355x.json"), "w") as f:
        json.dump({"fields": list(FIELDS), "plans": plans()}, f, indent=0)
    print("wrote plan_mi355x.json")
