"""Mixed-precision DIA and BSR parts against the FP64 oracle.  Error metric
per element: |y - y64| / max(|y64|, S) with S = sum_j |a_ij x_j| (plus |y0|):
bounded by 2e-5 for BF16 hi+lo (about 16 significant bits), 5e-4 for FP16
with a per-row scale (11 bits) and 0.08 for FP8 E4M3 with a per-slice scale
(4 bits: half an ulp is 1/16).  int16 delta columns give results bitwise
equal to int32 columns, and fall back to int32 when a step does not fit."""
import numpy as np
import pytest

from clspmv_amd import _lib as L
from clspmv_amd import matrix as M
from oracle import oracle as O
from tests.helpers import rand_coo

pytestmark = pytest.mark.gpu
BOUND = {L.VAL_BF16X2: 2e-5, L.VAL_FP16_ROW: 5e-4, L.VAL_FP8_SLICE: 0.08}
CODECS = [L.VAL_BF16X2, L.VAL_FP16_ROW, L.VAL_FP8_SLICE]


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


DIA_CASES = [("lap2d", lambda: M.gen_laplacian2d(64, 48)), ("stencil27", lambda: M.gen_stencil27(20, 3)),
             ("band_ragged", lambda: M.gen_stencil27(7, 4))]


@pytest.mark.parametrize("codec", CODECS)
@pytest.mark.parametrize("name,make", DIA_CASES, ids=[c[0] for c in DIA_CASES])
def test_dia_mixed(ctx, codec, name, make):
    coo = make()
    x, y0 = M.gen_vector(coo.width, 1), M.gen_vector(coo.height, 2)
    part = M.MixedPart(ctx, L.FMT_DIA, M.coo2dia(coo, 64), codec)
    assert part.info["method"] == 16 + codec
    e = _err(_run(ctx, part, coo, x, y0), coo, x, y0)
    assert e <= BOUND[codec], (name, codec, e)
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
    assert p16.info["method"] == 16 + codec + 4 and p32.info["method"] == 16 + codec
    ya = _run(ctx, p16, coo, x, y0)
    yb = _run(ctx, p32, coo, x, y0)
    assert np.array_equal(ya.view(np.uint32), yb.view(np.uint32))
    e = _err(ya, coo, x, y0)
    assert e <= BOUND[codec], (name, b, codec, e)
    p16.free()
    p32.free()


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
