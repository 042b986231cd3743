"""GPU parity at the benchmark sizes: each single-format matrix (27-point
stencil 128^3 as DIA, 4M x 32 uniform as ELL in the LDS x-tiled layout, 2M
Zipf(1.8) rows as CSR x-tiled, 1M block rows of 3x3 blocks as sliced BSR)
and the 4M-row Cocktail composite through the analyzer's MI355X plan, one
y = y0 + A x each, checked element by element against the oracle (the reference's FP32 CPU SpMV restated, with the FP64 product;
1e-5 relative bound, condition-aware: oracle.parity_report)."""
import numpy as np
import pytest

from clspmv_amd import cocktail as CK
from clspmv_amd import matrix as M
from clspmv_amd import workloads as W
from oracle import oracle as O

pytestmark = [pytest.mark.gpu, pytest.mark.slow]


def _once(ctx, part, x, y0):
    dx = M.DeviceArray.from_numpy(ctx, x)
    dy = M.DeviceArray.from_numpy(ctx, y0)
    part.spmv(dx, dy, 1)
    y = dy.download()
    dx.free()
    dy.free()
    return y


@pytest.mark.parametrize("name", ["c1a", "c1b", "c1c", "c1d"])
def test_single_format_full_size(ctx, name):
    wl = W.build(name)
    print(f"{name}: {wl.describe}, rows {wl.coo.height}, nnz {wl.coo.nnz}", flush=True)
    part = M.Part(ctx, wl.fmt, wl.mat, wl.method)
    wl.mat.free()
    x = M.gen_vector(wl.coo.width, 7)
    y0 = M.gen_vector(wl.coo.height, 8)
    y = _once(ctx, part, x, y0)
    kernel = part.info["kernel"]
    part.free()
    rep = O.parity_report(y, wl.coo, x, y0)
    wl.coo.free()
    assert rep["ok"], (name, kernel, rep)
    assert rep["max_err_f64"] < 1e-5, (name, kernel, rep)


def test_cocktail_composite_full_size(ctx):
    coo = W.make_coo("c2")
    print(f"c2: rows {coo.height}, nnz {coo.nnz}", flush=True)
    plan = CK.Plan.analyze(coo.copy(), CK.BenchTables(CK.MI355X_TABLES))
    x = M.gen_vector(coo.width, 7)
    y0 = M.gen_vector(coo.height, 8)
    gpu = CK.GpuCocktail(ctx, plan, x, y0)
    kernels = [p["kernel"] for p in gpu.part_infos()]
    assert len(kernels) == len(plan.parts()), kernels
    gpu.do_spmv(1)
    y = gpu.result_to_cpu()
    gpu.free()
    plan.free()
    rep = O.parity_report(y, coo, x, y0)
    coo.free()
    assert rep["ok"], (kernels, rep)
    assert rep["max_err_f64"] < 1e-5, (kernels, rep)
