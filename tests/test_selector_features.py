"""Structural features of the selector on matrices whose values are known in
closed form, the applicability rules, and the regret arithmetic.  CPU only."""
import math

import numpy as np

from clspmv_amd import matrix as M
from clspmv_amd import selector as S


def test_feature_count_and_names():
    assert len(S.FEATURE_NAMES) == 24
    assert len(S.features(M.gen_laplacian2d(16, 16))) == 24


def test_laplacian_features_closed_form():
    n = 64
    f = dict(zip(S.FEATURE_NAMES, S.features(M.gen_laplacian2d(n, n))))
    rows = n * n
    nnz = 5 * rows - 4 * n
    assert f["log2_rows"] == math.log2(rows) and f["log2_nnz"] == math.log2(nnz)
    assert abs(f["mean_row"] - nnz / rows) < 1e-12
    assert f["nnz_top7_diagonals"] == 1.0 and f["log2_diagonals"] == math.log2(6)
    assert abs(f["log2_dia_padding"] - math.log2(5 * rows / nnz)) < 1e-12
    assert f["empty_rows"] == 0.0 and f["rows_over_64"] == 0.0
    # interior row: (r-n, r-1, r, r+1, r+n): two adjacent pairs per 3-run
    assert 0.35 < f["adjacent_column_pairs"] < 0.41


def test_uniform_rows_and_blocks():
    f = dict(zip(S.FEATURE_NAMES, S.features(M.gen_uniform_rows(4096, 16, 1))))
    assert f["mean_row"] == 16 and f["row_cv"] == 0.0 and f["log2_ell_padding"] == 0.0
    g = dict(zip(S.FEATURE_NAMES, S.features(M.gen_block_rows(1000, 4, 6, 100, 2))))
    assert g["block_fill_4"] == 1.0 and g["block_fill_2"] == 1.0


def test_applicability():
    lap = S.features(M.gen_laplacian2d(64, 64))
    rnd = S.features(M.gen_uniform_rows(4096, 8, 3))
    ok = S.applicable(np.stack([lap, rnd]))
    names = S.CANDIDATE_NAMES
    assert ok[0, names.index("dia_row")] and not ok[1, names.index("dia_row")]
    assert ok[1, names.index("ell_row4")] and ok[1, names.index("csr_adaptive")]
    assert not ok[1, names.index("bsr3_sliced")]


def test_regret():
    t = np.array([[1.0, 2.0, np.inf], [3.0, 1.5, 1.0]])
    assert np.allclose(S.regret(t, np.array([0, 1])), [0.0, 0.5])
