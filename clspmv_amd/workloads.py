"""Seeded benchmark workloads (synthetic matrices) and their format choice.

Configurations (all square, values N(0,1) unless stated, no duplicate
(row, col)):
  c0   2D 5-point Laplacian on 256 x 256 (65,536 rows, 326,656 nnz), 4 on the
       diagonal and -1 off it; x ~ U[-1, 1]
  c1a  3D 27-point stencil on 128^3 (2,097,152 rows, 382^3 nnz)      -> DIA
  c1b  4,194,304 rows x exactly 32 uniform distinct columns           -> ELL
  c1c  2,097,152 rows, Zipf(1.8) row lengths rescaled to mean 16,
       max 50,000 (one row forced to 50,000), uniform columns         -> CSR
  c1d  1,048,576 block rows of dense 3x3 blocks, 9 distinct block
       columns per block row within +-5,000 of the diagonal           -> BSR 3x3
  c2   Cocktail composite, 4,194,304 rows (band + blocks + Zipf(1.6))
  c4   batch: matrices drawn from the c1 generators with row counts
       log-uniform in [2^10, 2^20], each with k in {1, 4, 16} right-hand sides
Sizes can be scaled down (``scale``) for tests; the seeds are fixed.
"""
from __future__ import annotations

from dataclasses import dataclass

from . import _lib as L
from . import matrix as M

SEEDS = {"c0": 0, "c1a": 101, "c1b": 102, "c1c": 103, "c1d": 104, "c2": 200}


@dataclass
class Workload:
    name: str
    coo: M.CooMatrix
    fmt: int
    method: int
    mat: object  # converted format (owns its arrays)
    describe: str


def make_coo(name: str, scale: int = 1):
    """Generate the COO of one configuration; `scale` divides the row count."""
    s = SEEDS[name]
    if name == "c0":
        return M.gen_laplacian2d(256 // scale if scale > 1 else 256, 256 // scale if scale > 1 else 256)
    if name == "c1a":
        n = 128 // round(scale ** (1 / 3)) if scale > 1 else 128
        return M.gen_stencil27(n, s)
    if name == "c1b":
        return M.gen_uniform_rows(4194304 // scale, 32, s)
    if name == "c1c":
        rows = 2097152 // scale
        return M.gen_zipf_rows(rows, 1.8, 16.0, min(50000, rows // 2), s)
    if name == "c1d":
        return M.gen_block_rows(1048576 // scale, 3, 9, 5000, s)
    if name == "c2":
        return M.gen_cocktail_composite(4194304 // scale, s)
    raise KeyError(name)


def best_format(name: str, coo):
    """The single format each c1 generator is meant to exercise."""
    if name in ("c0", "c1a"):
        return L.FMT_DIA, 0, M.coo2dia(coo), "DIA"
    if name == "c1b":
        return L.FMT_ELL, 4, M.coo2ell(coo), "ELL (LDS x-tiled)"
    if name == "c1c":
        return L.FMT_CSR_VECTOR, 4, M.coo2csr(coo), "CSR (LDS x-tiled)"
    if name == "c1d":
        return L.FMT_BSR, 0, M.coo2bsr(coo, 3, 3), "BSR 3x3 (sliced)"
    raise KeyError(name)


def reference_bytes(name: str, coo) -> int:
    """Algorithmic bytes of one y += A x in the reference format each c1
    matrix is named for: the format's stored arrays (values, indices,
    pointers, its own padding) + 4 B per column (x read once) + 8 B per row
    (y read and written).  DIA: 4 B per slot of ndiag x align(rows, 32);
    ELL: 8 B per slot of max-row-length x align(rows, 32); CSR: 8 B per
    nonzero + 4 B per row pointer; BCSR b x b: 4 b^2 + 4 B per block + 4 B
    per block-row pointer."""
    import numpy as np

    rows, cols, nnz = coo.height, coo.width, coo.nnz
    xy = 4 * cols + 8 * rows
    al = (rows + 31) // 32 * 32
    if name in ("c0", "c1a"):
        ndiag = np.unique(coo.col_id.astype(np.int64) - coo.row_id).size
        return 4 * ndiag * al + xy
    if name == "c1b":
        width = int(np.bincount(coo.row_id, minlength=rows).max()) if nnz else 0
        return 8 * width * al + xy
    if name == "c1c":
        return 8 * nnz + 4 * (rows + 1) + xy
    if name == "c1d":
        b = 3
        keys = np.unique((coo.row_id.astype(np.int64) // b) * ((cols + b - 1) // b) + coo.col_id // b)
        return (4 * b * b + 4) * keys.size + 4 * ((rows + b - 1) // b + 1) + xy
    raise KeyError(name)


def build(name: str, scale: int = 1) -> Workload:
    coo = make_coo(name, scale)
    fmt, meth, mat, desc = best_format(name, coo)
    return Workload(name, coo, fmt, meth, mat, desc)


def c4_matrices(n: int = 1024, seed: int = 400, log2_min: float = 10, log2_max: float = 20):
    """The batch workload: (kind, coo, k) for n seeded matrices from the c1
    generators (27-point stencil, uniform 32 per row, Zipf(1.8) mean 16,
    3x3 blocks x 9) at log-uniform row counts, k cycling 1, 4, 16."""
    import numpy as np

    rng = np.random.default_rng(seed)
    kinds = ["stencil27", "uniform", "zipf", "blocks"]
    out = []
    for i in range(n):
        kind = kinds[i % 4]
        rows = int(2 ** rng.uniform(log2_min, log2_max))
        s = seed * 100003 + i
        if kind == "stencil27":
            coo = M.gen_stencil27(max(4, round(rows ** (1 / 3))), s)
        elif kind == "uniform":
            coo = M.gen_uniform_rows(rows, 32, s)
        elif kind == "zipf":
            coo = M.gen_zipf_rows(rows, 1.8, 16.0, min(50000, rows // 2), s)
        else:
            coo = M.gen_block_rows(max(16, rows // 3), 3, 9, 5000, s)
        out.append((kind, coo, (1, 4, 16)[i % 3]))
    return out
