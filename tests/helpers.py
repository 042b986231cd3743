"""Shared helpers of the test-suite (inputs, GPU runs through the C ABI)."""
import numpy as np

from clspmv_amd import matrix as M


def rand_coo(h, w, nnz, seed):
    """Random sorted COO with distinct (row, col) and N(0,1) values."""
    rng = np.random.default_rng(seed)
    nnz = min(nnz, h * w)
    keys = rng.choice(h * w, size=nnz, replace=False) if nnz else np.zeros(0, np.int64)
    r = (keys // w).astype(np.int32)
    c = (keys % w).astype(np.int32)
    v = rng.standard_normal(r.size).astype(np.float32)
    o = np.lexsort((c, r))
    return M.CooMatrix.from_arrays(r[o], c[o], v[o], w, h)


def gpu_spmv(ctx, fmt, mat, x, y0, method=0, ntimes=1):
    """Upload `mat` as a part, run ntimes * (y += A x), return y and info."""
    part = M.Part(ctx, fmt, mat, method)
    dx = M.DeviceArray.from_numpy(ctx, np.asarray(x, np.float32))
    dy = M.DeviceArray.from_numpy(ctx, np.asarray(y0, np.float32))
    part.spmv(dx, dy, ntimes)
    y = dy.download()
    info = dict(part.info)
    part.free()
    dx.free()
    dy.free()
    return y, info
