"""Inputs of the analyzer parity fixtures: seeded generator calls (all
deterministic) and the analyzer knob settings
(dia_priority_multiplier, estimate_flat_partial, estimate_flat_full)."""
import hashlib

import numpy as np

from clspmv_amd import matrix as M

KNOBS = [(1.0, 1, 0), (0.5, 1, 0), (1.0, 0, 0), (1.0, 1, 1)]

CASES = [
    ("lap2d_64", ["lap2d", 64]),
    ("lap2d_256", ["lap2d", 256]),
    ("st27_24", ["st27", 24, 1]),
    ("st27_40", ["st27", 40, 12]),
    ("uni_20k_32", ["uniform", 20000, 32, 2]),
    ("zipf_20k", ["zipf", 20000, 1.8, 16.0, 5000, 3]),
    ("blk3", ["blocks", 8000, 3, 9, 300, 4]),
    ("blk4", ["blocks", 8000, 4, 6, 300, 5]),
    ("cocktail_64k", ["cocktail", 65536, 6]),
    ("band15_rand8", ["band_rand", 40000, -7, 7, 8, 3]),
    ("band7_rand2", ["band_rand", 40000, -3, 3, 2, 2]),
    ("blk4_rand", ["blocks_rand", 10000, 4, 6, 200, 3, 6]),
    ("rand_ragged", ["rand", 3001, 2999, 40000, 7]),
    ("rand_rect", ["rand", 500, 3000, 20000, 8]),
]

MAXFLOP_POINTS = [(1, 0), (7, 20), (1000, 5000), (65536, 326656), (131072, 131072 * 27), (2097152, 55742968),
                  (4194304, 134217728), (2097152, 33554432), (3145728, 84934656), (4194304, 59800000),
                  (300000, 30000000), (12345, 678901), (5000000, 1000)]


def _rand(h, w, nnz, seed):
    rng = np.random.default_rng(seed)
    keys = rng.choice(h * w, size=min(nnz, h * w), replace=False)
    r = (keys // w).astype(np.int32)
    c = (keys % w).astype(np.int32)
    v = rng.standard_normal(r.size).astype(np.float32)
    return r, c, v


def _combine(n, parts):
    r = np.concatenate([p[0] for p in parts])
    c = np.concatenate([p[1] for p in parts])
    v = np.concatenate([p[2] for p in parts])
    key = r.astype(np.int64) * n + c
    _, idx = np.unique(key, return_index=True)
    o = np.lexsort((c[idx], r[idx]))
    idx = idx[o]
    return M.CooMatrix.from_arrays(r[idx], c[idx], v[idx], n, n)


def make(gen):
    kind = gen[0]
    if kind == "lap2d":
        return M.gen_laplacian2d(gen[1], gen[1])
    if kind == "st27":
        return M.gen_stencil27(gen[1], gen[2])
    if kind == "uniform":
        return M.gen_uniform_rows(*gen[1:])
    if kind == "zipf":
        return M.gen_zipf_rows(*gen[1:])
    if kind == "blocks":
        return M.gen_block_rows(*gen[1:])
    if kind == "cocktail":
        return M.gen_cocktail_composite(*gen[1:])
    if kind == "rand":
        h, w, nnz, seed = gen[1:]
        r, c, v = _rand(h, w, nnz, seed)
        o = np.lexsort((c, r))
        return M.CooMatrix.from_arrays(r[o], c[o], v[o], w, h)
    if kind == "band_rand":
        n, lo, hi, per_row, seed = gen[1:]
        rng = np.random.default_rng(seed)
        rr, cc = [], []
        for o in range(lo, hi + 1):
            rows = np.arange(max(0, -o), min(n, n - o))
            rr.append(rows)
            cc.append(rows + o)
        r = np.concatenate(rr).astype(np.int32)
        c = np.concatenate(cc).astype(np.int32)
        band = (r, c, rng.standard_normal(r.size).astype(np.float32))
        return _combine(n, [band, _rand(n, n, per_row * n, seed + 1)])
    if kind == "blocks_rand":
        brows, b, per, bandw, per_row, seed = gen[1:]
        blk = M.gen_block_rows(brows, b, per, bandw, seed)
        n = brows * b
        return _combine(n, [(blk.row_id.copy(), blk.col_id.copy(), blk.data.copy()), _rand(n, n, per_row * n, seed + 1)])
    raise KeyError(kind)


def input_digest(coo):
    h = hashlib.sha256()
    for a in (coo.row_id, coo.col_id, coo.data):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()
