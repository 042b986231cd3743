"""Writes tests/golden/analyzer.json: the reference's Cocktail partitions
(analyze_matrix compiled from the reference sources into oracle/_ref) for
seeded inputs and four settings of the analyzer knobs, plus the reference's
maxflop_info on a grid of (rows, nnz) points.

    python tests/golden/make_golden_analyzer.py

Each partition is stored as its scalar fields and the SHA-256 of every
defined array; the input is identified by its generator call and the SHA-256
of its (row, col, value) arrays, so a fixture also detects a changed input.
"""
from __future__ import annotations

import ctypes as C
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from clspmv_amd import _lib as L  # noqa: E402
from oracle import oracle as O  # noqa: E402
from tests import analyzer_cases as AC  # noqa: E402
from tests.helpers import cocktail_arrays  # noqa: E402

REF_TREE = "/root/reference/clSpMV_AMD"


def digest(d):
    out = {}
    for k, v in d.items():
        if isinstance(v, np.ndarray):
            out[k] = hashlib.sha256(np.ascontiguousarray(v).tobytes()).hexdigest()
        else:
            out[k] = list(v) if isinstance(v, tuple) else v
    return out


def main():
    os.environ["CLSPMVPATH"] = REF_TREE
    R = O.ref()
    if R is None:
        sys.exit("oracle/_ref/libclspmv_ref.so is not built")
    R.ref_analyze_matrix.argtypes = [C.c_void_p, C.c_void_p, C.c_float, C.c_int, C.c_int]
    R.ref_analyze_matrix.restype = None
    R.ref_maxflop.argtypes = [C.c_int, C.c_int, C.c_void_p]
    R.ref_maxflop.restype = None
    lib = L.lib()
    # the reference prints its decisions on stdout: keep them out of the way
    devnull = os.open(os.devnull, os.O_WRONLY)
    saved = os.dup(1)
    os.dup2(devnull, 1)
    try:
        out = {"partitions": [], "maxflop": []}
        for name, gen in AC.CASES:
            coo = AC.make(gen)
            for knobs in AC.KNOBS:
                r = L.Cocktail()
                lib.clspmv_init_cocktail(C.byref(r))
                c2 = coo.copy()
                R.ref_analyze_matrix(C.addressof(c2.s), C.addressof(r), knobs[0], knobs[1], knobs[2])
                out["partitions"].append({"case": name, "gen": gen, "input_sha256": AC.input_digest(coo),
                                          "knobs": list(knobs), "result": digest(cocktail_arrays(r))})
        for rows, nnz in AC.MAXFLOP_POINTS:
            m = L.MaxflopInfo()
            R.ref_maxflop(rows, nnz, C.addressof(m))
            rec = {}
            for f, _ in L.MaxflopInfo._fields_:
                v = getattr(m, f)
                rec[f] = list(v) if hasattr(v, "__len__") else v
            if rec["csr_max_flop"] == 0:
                rec.pop("csr_groupnum")
            if rec["coo_max_flop"] == 0:
                rec.pop("coo_groupnum")
            out["maxflop"].append({"rows": rows, "nnz": nnz, "info": rec})
    finally:
        os.dup2(saved, 1)
    with open(os.path.join(HERE, "analyzer.json"), "w") as f:
        json.dump(out, f, indent=0)
    print("wrote analyzer.json with", len(out["partitions"]), "partitions")


if __name__ == "__main__":
    main()
