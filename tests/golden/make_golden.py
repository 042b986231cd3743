"""Writes the golden fixtures of tests/golden from the reference's own
compiled CPU code (oracle/_ref/libclspmv_ref.so, built by oracle/Makefile
from the reference sources in this container).

    python tests/golden/make_golden.py

Outputs (data only: inputs and the reference's outputs):
  conversions.npz   seeded small COO inputs and, for each, every array the
                    reference's coo2csr / coo2dia / coo2bdia / coo2ell /
                    coo2sell(32,128) / coo2b4csr / coo2b4ell / coo2sbell
                    (8 block shapes) / pad_csr(4,16) produce, plus its
                    spmv_only result for a seeded x
  mmf/*.mtx + mmf.npz   small Matrix Market files (real, symmetric, pattern,
                    integer, comments) and the COO the reference's ReadMMF
                    reads from each
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from clspmv_amd import _lib as L  # noqa: E402
from oracle import oracle as O  # noqa: E402

SHAPES = [(bh, bw) for bw in (4, 8) for bh in (1, 2, 4, 8)]


def arr(p, n, dt):
    return np.ctypeslib.as_array(p, shape=(int(n),)).copy() if n > 0 else np.zeros(0, dt)


def make_inputs():
    """Small seeded COO inputs covering ragged shapes, empty rows/columns,
    a stencil, long rows and a non-square matrix."""
    rng = np.random.default_rng(20240601)
    cases = []
    for i, (h, w, nnz) in enumerate([(7, 7, 12), (33, 33, 200), (64, 50, 300), (50, 64, 260), (97, 97, 97),
                                     (130, 130, 2000), (1, 9, 5), (40, 40, 0), (200, 180, 900)]):
        nnz = min(nnz, h * w)
        keys = rng.choice(h * w, size=nnz, replace=False) if nnz else np.zeros(0, np.int64)
        r, c = (keys // w).astype(np.int32), (keys % w).astype(np.int32)
        o = np.lexsort((c, r))
        v = rng.standard_normal(nnz).astype(np.float32)
        cases.append((f"rand{i}_{h}x{w}", r[o], c[o], v, w, h))
    # 2D 5-point Laplacian 12x12 (closed-form structure)
    n = 12
    rr, cc, vv = [], [], []
    for row in range(n * n):
        ix, iy = row % n, row // n
        for dc, val in ((-n, -1.0), (-1, -1.0), (0, 4.0), (1, -1.0), (n, -1.0)):
            col = row + dc
            if dc == -n and iy == 0 or dc == n and iy == n - 1 or dc == -1 and ix == 0 or dc == 1 and ix == n - 1:
                continue
            rr.append(row)
            cc.append(col)
            vv.append(val)
    cases.append(("lap2d_12", np.array(rr, np.int32), np.array(cc, np.int32), np.array(vv, np.float32), n * n, n * n))
    # long rows next to short ones
    h, w = 20, 300
    rr, cc = [], []
    for row in range(h):
        l = 250 if row in (3, 4, 15) else int(rng.integers(0, 6))
        cols = np.sort(rng.choice(w, size=l, replace=False))
        rr += [row] * l
        cc += list(cols)
    cases.append(("longrows_20x300", np.array(rr, np.int32), np.array(cc, np.int32),
                  rng.standard_normal(len(rr)).astype(np.float32), w, h))
    return cases


def coo_struct(r, c, v, w, h):
    s = L.Coo()
    L.check(L.lib().clspmv_alloc_coo(C.byref(s), int(w), int(h), int(r.size)), "alloc")
    if r.size:
        arr_r = np.ctypeslib.as_array(s.coo_row_id, shape=(r.size,))
        arr_c = np.ctypeslib.as_array(s.coo_col_id, shape=(r.size,))
        arr_v = np.ctypeslib.as_array(s.coo_data, shape=(r.size,))
        arr_r[:], arr_c[:], arr_v[:] = r, c, v
    return s


def conversions(R, out):
    for name, r, c, v, w, h in make_inputs():
        key = name
        out[f"{key}/in_row"], out[f"{key}/in_col"], out[f"{key}/in_val"] = r, c, v
        out[f"{key}/shape"] = np.array([h, w], np.int32)
        # CSR
        s = coo_struct(r, c, v, w, h)
        o = L.Csr()
        R.ref_coo2csr(C.addressof(s), C.addressof(o))
        out[f"{key}/csr_row_ptr"] = arr(o.csr_row_ptr, h + 1, np.int32)
        # pad_csr (the reference reads padcol[-1] when the first row is
        # empty: undefined behaviour, no fixture for such inputs)
        for a in ((4, 16) if h and out[f"{key}/csr_row_ptr"][1] > 0 else ()):
            p = L.Csr()
            R.ref_pad_csr(C.addressof(o), C.addressof(p), a)
            out[f"{key}/padcsr{a}_row_ptr"] = arr(p.csr_row_ptr, h + 1, np.int32)
            out[f"{key}/padcsr{a}_col"] = arr(p.csr_col_id, p.matinfo.nnz, np.int32)
            out[f"{key}/padcsr{a}_val"] = arr(p.csr_data, p.matinfo.nnz, np.float32)
        # DIA / BDIA (the reference's dia needs rows < min(w, h) aligned)
        if h <= w or True:
            s = coo_struct(r, c, v, w, h)
            d = L.Dia()
            ok = R.ref_coo2dia(C.addressof(s), C.addressof(d), 32) if h <= w else -1
            out[f"{key}/dia_ok"] = np.array([ok], np.int32)
            if ok == 1 and d.dia_num == 0:
                ok = -1  # coo2bdia reads offsets[0] of an empty array: no fixture
                out[f"{key}/dia_ok"] = np.array([ok], np.int32)
            if ok == 1:
                out[f"{key}/dia_offsets"] = arr(d.dia_offsets, d.dia_num, np.int32)
                out[f"{key}/dia_data"] = arr(d.dia_data, d.dia_num * d.dia_length_aligned, np.float32)
                out[f"{key}/dia_meta"] = np.array([d.dia_num, d.dia_length, d.dia_length_aligned], np.int32)
                s = coo_struct(r, c, v, w, h)
                b = L.Bdia()
                R.ref_coo2bdia(C.addressof(s), C.addressof(b), 32)
                out[f"{key}/bdia_offsets"] = arr(b.bdia_offsets, b.bdia_band_num, np.int32)
                out[f"{key}/bdia_bptr"] = arr(b.bdia_bptr, b.bdia_band_num + 1, np.int32)
            elif ok == 0:
                out[f"{key}/dia_meta"] = np.array([d.dia_num, d.dia_length, d.dia_length_aligned], np.int32)
        # ELL
        s = coo_struct(r, c, v, w, h)
        e = L.Ell()
        R.ref_coo2ell(C.addressof(s), C.addressof(e), 32, 0)
        n = e.ell_num * e.ell_height_aligned
        out[f"{key}/ell_meta"] = np.array([e.ell_num, e.ell_height_aligned], np.int32)
        out[f"{key}/ell_col"] = arr(e.ell_col_id, n, np.int32)
        out[f"{key}/ell_val"] = arr(e.ell_data, n, np.float32)
        # SELL
        for S in (32, 128):
            s = coo_struct(r, c, v, w, h)
            sl = L.Sell()
            R.ref_coo2sell(C.addressof(s), C.addressof(sl), S)
            sp = arr(sl.sell_slice_ptr, sl.sell_slice_num + 1, np.int32)
            out[f"{key}/sell{S}_ptr"] = sp
            out[f"{key}/sell{S}_col"] = arr(sl.sell_col_id, sp[-1], np.int32)
            out[f"{key}/sell{S}_val"] = arr(sl.sell_data, sp[-1], np.float32)
        # blocked formats
        for bh, bw in SHAPES:
            tag = f"{bh}x{bw}"
            s = coo_struct(r, c, v, w, h)
            b4 = L.B4Csr()
            ok = R.ref_coo2b4csr(C.addressof(s), C.addressof(b4), bw, bh, 32)
            out[f"{key}/b4csr{tag}_row_ptr"] = arr(b4.b4csr_row_ptr, b4.b4csr_row_num + 1, np.int32)
            out[f"{key}/b4csr{tag}_col"] = arr(b4.b4csr_col_id, b4.b4csr_block_num, np.int32)
            out[f"{key}/b4csr{tag}_val"] = arr(b4.b4csr_data, b4.b4csr_aligned_size * bh * (bw // 4), np.float32)
            s = coo_struct(r, c, v, w, h)
            be = L.B4Ell()
            R.ref_coo2b4ell(C.addressof(s), C.addressof(be), bw, bh, 32, 0)
            brows, Hc, K, F = be.b4ell_row_num, be.b4ell_height_aligned, be.b4ell_block_num, be.b4ell_float4_aligned
            # only the entries of real block rows are defined by the reference
            colm = arr(be.b4ell_col_id, Hc * K, np.int32).reshape(K, Hc)[:, :brows]
            datm = arr(be.b4ell_data, F * bh * (bw // 4) * K, np.float32).reshape(K * (bw // 4) * bh, F)[:, :4 * brows]
            out[f"{key}/b4ell{tag}_meta"] = np.array([brows, Hc, K, F], np.int32)
            out[f"{key}/b4ell{tag}_col"] = np.ascontiguousarray(colm)
            out[f"{key}/b4ell{tag}_val"] = np.ascontiguousarray(datm)
            s = coo_struct(r, c, v, w, h)
            sb = L.Sbell()
            R.ref_coo2sbell(C.addressof(s), C.addressof(sb), bw, bh, 32)
            sp = arr(sb.sbell_slice_ptr, sb.sbell_slice_num + 1, np.int32)
            out[f"{key}/sbell{tag}_ptr"] = sp
            out[f"{key}/sbell{tag}_col"] = arr(sb.sbell_col_id, sp[-1], np.int32)
            out[f"{key}/sbell{tag}_val"] = arr(sb.sbell_data, sp[-1] * bw * bh, np.float32)
        # spmv_only
        x = np.random.default_rng(7).uniform(-1, 1, w).astype(np.float32)
        y = np.zeros(h, np.float32)
        s = coo_struct(r, c, v, w, h)
        R.ref_spmv_only(C.addressof(s), x.ctypes.data_as(C.POINTER(C.c_float)), y.ctypes.data_as(C.POINTER(C.c_float)))
        out[f"{key}/spmv_x"] = x
        out[f"{key}/spmv_y"] = y


MMF = {
    "real_general.mtx": "%%MatrixMarket matrix coordinate real general\n% comment line\n%another\n4 5 6\n"
                        "1 1 1.5\n2 3 -2.25\n4 5 3e-1\n3 2 7\n1 4 0.1\n4 1 -1e3\n",
    "real_symmetric.mtx": "%%MatrixMarket matrix coordinate real symmetric\n5 5 7\n1 1 2.0\n2 1 -1.0\n"
                          "3 1 0.5\n3 3 4.0\n5 2 1.25\n4 4 -3.0\n5 5 9.5\n",
    "pattern_general.mtx": "%%MatrixMarket matrix coordinate pattern general\n3 3 4\n1 2\n2 3\n3 1\n2 2\n",
    "pattern_symmetric.mtx": "%%MatrixMarket matrix coordinate pattern symmetric\n4 4 4\n2 1\n3 3\n4 2\n4 1\n",
    # "integer" is not "real": the reference reads no value token, so each value
    # becomes the next row index (a quirk the reader must reproduce)
    "integer_general.mtx": "%%MatrixMarket matrix coordinate integer general\n6 6 2\n1 1 3\n2 2 4\n",
    "unsorted_dup.mtx": "%%MatrixMarket matrix coordinate real general\n3 4 6\n3 1 1\n1 4 2\n2 2 3\n1 1 4\n"
                        "3 1 5\n2 2 6\n",
}


def mmf(R, out):
    d = os.path.join(HERE, "mmf")
    os.makedirs(d, exist_ok=True)
    for fname, text in MMF.items():
        path = os.path.join(d, fname)
        with open(path, "w") as f:
            f.write(text)
        s = L.Coo()
        R.ref_read_mmf(path.encode(), C.addressof(s))
        n = s.matinfo.nnz
        out[f"{fname}/shape"] = np.array([s.matinfo.height, s.matinfo.width, n], np.int32)
        out[f"{fname}/row"] = arr(s.coo_row_id, n, np.int32)
        out[f"{fname}/col"] = arr(s.coo_col_id, n, np.int32)
        out[f"{fname}/val"] = arr(s.coo_data, n, np.float32)


def main():
    R = O.ref()
    if R is None:
        sys.exit("oracle/_ref/libclspmv_ref.so is not built (it needs the reference tree)")
    out = {}
    conversions(R, out)
    np.savez_compressed(os.path.join(HERE, "conversions.npz"), **out)
    out = {}
    mmf(R, out)
    np.savez_compressed(os.path.join(HERE, "mmf.npz"), **out)
    print("wrote", os.listdir(HERE))


if __name__ == "__main__":
    main()
