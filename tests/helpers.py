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


# ---------------------------------------------------------------- Cocktail
import ctypes as _C  # noqa: E402

from clspmv_amd import _lib as _L  # noqa: E402


def _arr(p, n, dt):
    return np.ctypeslib.as_array(p, shape=(int(n),)).copy() if n > 0 and p else np.zeros(0, dt)


def cocktail_arrays(c):
    """Every field and defined array of a Cocktail struct as a dict (the
    entries the reference leaves uninitialised are excluded: b4ell columns of
    block rows past the last, sbell columns of block rows past the last)."""
    out = {}
    for n in ("bdia", "dia", "sbell", "bell", "bcsr", "sell", "ell", "csr", "coo"):
        out[f"ifuse{n}"] = int(getattr(c, f"ifuse{n}"))
        if getattr(c, f"ifuse{n}"):
            out[f"{n}_meth_num"] = getattr(c, f"{n}_meth_num")
    if c.ifusecsr:
        out["csr_group_num"] = c.csr_group_num
    if c.ifusecoo:
        out["coo_group_num"] = c.coo_group_num
    out["shape"] = (c.mat_width, c.mat_height, c.mat_nnz)
    if c.ifusebdia:
        b = c.bdia
        out["bdia.meta"] = (b.matinfo.nnz, b.bdia_band_num, b.bdia_length, b.bdia_length_aligned)
        out["bdia.offsets"] = _arr(b.bdia_offsets, b.bdia_band_num, np.int32)
        bp = _arr(b.bdia_bptr, b.bdia_band_num + 1, np.int32)
        out["bdia.bptr"] = bp
        out["bdia.data"] = _arr(b.bdia_data, int(bp[-1]) * b.bdia_length_aligned, np.float32)
    if c.ifusedia:
        d = c.dia
        out["dia.meta"] = (d.matinfo.nnz, d.dia_num, d.dia_length, d.dia_length_aligned)
        out["dia.offsets"] = _arr(d.dia_offsets, d.dia_num, np.int32)
        out["dia.data"] = _arr(d.dia_data, d.dia_num * d.dia_length_aligned, np.float32)
    if c.ifusesbell:
        s = c.sbell
        out["sbell.meta"] = (s.matinfo.nnz, s.sbell_bwidth, s.sbell_bheight, s.sbell_slice_height, s.sbell_slice_num,
                             s.sbell_row_num)
        sp = _arr(s.sbell_slice_ptr, s.sbell_slice_num + 1, np.int32)
        out["sbell.ptr"] = sp
        col = _arr(s.sbell_col_id, sp[-1], np.int32)
        # only block rows < sbell_row_num are defined
        keep = np.zeros(col.size, bool)
        S = s.sbell_slice_height
        for sl in range(s.sbell_slice_num):
            w = (sp[sl + 1] - sp[sl]) // S
            for t in range(S):
                if sl * S + t < s.sbell_row_num:
                    keep[sp[sl] + t + S * np.arange(w)] = True
        out["sbell.col"] = col[keep]
        out["sbell.data"] = _arr(s.sbell_data, sp[-1] * s.sbell_bwidth * s.sbell_bheight, np.float32)
    if c.ifusebell:
        b = c.bell
        out["bell.meta"] = (b.matinfo.nnz, b.b4ell_bwidth, b.b4ell_bheight, b.b4ell_row_num, b.b4ell_block_num,
                            b.b4ell_height_aligned, b.b4ell_float4_aligned)
        K, Hc, F = b.b4ell_block_num, b.b4ell_height_aligned, b.b4ell_float4_aligned
        out["bell.col"] = _arr(b.b4ell_col_id, Hc * K, np.int32).reshape(K, Hc)[:, :b.b4ell_row_num].copy()
        out["bell.data"] = _arr(b.b4ell_data, F * b.b4ell_bheight * (b.b4ell_bwidth // 4) * K, np.float32)
    if c.ifusebcsr:
        b = c.bcsr
        out["bcsr.meta"] = (b.matinfo.nnz, b.b4csr_bwidth, b.b4csr_bheight, b.b4csr_row_num, b.b4csr_aligned_size,
                            b.b4csr_block_num)
        out["bcsr.row_ptr"] = _arr(b.b4csr_row_ptr, b.b4csr_row_num + 1, np.int32)
        out["bcsr.col"] = _arr(b.b4csr_col_id, b.b4csr_block_num, np.int32)
        out["bcsr.data"] = _arr(b.b4csr_data, b.b4csr_aligned_size * b.b4csr_bheight * (b.b4csr_bwidth // 4),
                                np.float32)
    if c.ifusesell:
        s = c.sell
        out["sell.meta"] = (s.matinfo.nnz, s.sell_slice_height, s.sell_slice_num)
        sp = _arr(s.sell_slice_ptr, s.sell_slice_num + 1, np.int32)
        out["sell.ptr"] = sp
        out["sell.col"] = _arr(s.sell_col_id, sp[-1], np.int32)
        out["sell.data"] = _arr(s.sell_data, sp[-1], np.float32)
    if c.ifuseell:
        e = c.ell
        out["ell.meta"] = (e.matinfo.nnz, e.ell_num, e.ell_height_aligned)
        out["ell.col"] = _arr(e.ell_col_id, e.ell_num * e.ell_height_aligned, np.int32)
        out["ell.data"] = _arr(e.ell_data, e.ell_num * e.ell_height_aligned, np.float32)
    if c.ifusecsr:
        m = c.csr
        out["csr.meta"] = (m.matinfo.nnz,)
        out["csr.row_ptr"] = _arr(m.csr_row_ptr, m.matinfo.height + 1, np.int32)
        out["csr.col"] = _arr(m.csr_col_id, m.matinfo.nnz, np.int32)
        out["csr.data"] = _arr(m.csr_data, m.matinfo.nnz, np.float32)
    if c.ifusecoo:
        m = c.coo
        out["coo.meta"] = (m.matinfo.nnz,)
        out["coo.row"] = _arr(m.coo_row_id, m.matinfo.nnz, np.int32)
        out["coo.col"] = _arr(m.coo_col_id, m.matinfo.nnz, np.int32)
        out["coo.data"] = _arr(m.coo_data, m.matinfo.nnz, np.float32)
    return out


def cocktail_diff(a, b):
    """Names of the fields that differ between two cocktail_arrays dicts."""
    bad = sorted(set(a) ^ set(b))
    for k in set(a) & set(b):
        x, y = a[k], b[k]
        if isinstance(x, np.ndarray):
            if x.shape != y.shape or not np.array_equal(x.view(np.uint32) if x.dtype == np.float32 else x,
                                                        y.view(np.uint32) if y.dtype == np.float32 else y):
                bad.append(k)
        elif x != y:
            bad.append(k)
    return bad
