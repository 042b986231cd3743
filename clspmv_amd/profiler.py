"""MI355X performance tables for the Cocktail analyzer.

The analyzer picks formats from ten per-format tables of measured GFLOP/s
(A/src/analyze.cpp:16-210 reads them; A/src/bench_*.cpp wrote the shipped
ones on an AMD GPU).  This module measures this library's kernels on
the same family of synthetic matrices -- square, banded, every row with the
same number of entries (or block rows with the same number of dense blocks),
values 1.0 -- and writes tables with the same schema, so that the analyzer
(unchanged, bit-exact with the reference) makes MI355X choices:

    overhead.ben  groups threads usec
    dia.ben       dim diagonals  gflops[4]      (row, 4-rows, -, -)
    bdia.ben      dim diagonals  gflops[5]      (direct x, LDS x band, -, -, -)
    ell.ben       dim ell_num    gflops[2]      (row, 4-rows)
    sell.ben      dim width      gflops[2]      (slice 32, slice 128)
    csr.ben       dim width groups gflops[2]    (adaptive row blocks, -)
    coo.ben       dim width groups gflops[1]    (segmented reduction)
    bell.ben      bh bw dim blocks gflops[2]    (b4ell kernel, -)
    sbell.ben     bh bw dim blocks gflops[1]
    bcsr.ben      bh bw dim blocks gflops[1]

A "-" column is a reference kernel variant with no counterpart here
(texture reads, workgroup-count variants); it is written as 0 so the
analyzer never selects it, and the Cocktail driver maps every method number
onto the kernel measured in the column it came from.  The CSR/COO kernels do
not take a workgroup count, so their tables carry one group entry (the CU
count).  Records whose stored arrays would exceed a size cap are skipped
(the analyzer clamps to the largest width it has for a dimension).

    python -m clspmv_amd.profiler --out clspmv_amd/tables/mi355x [--quick]
"""
from __future__ import annotations

import argparse
import ctypes as C
import datetime
import os
import sys
import time

import numpy as np

from . import _lib as L
from . import matrix as M
from ._lib import check, lib

DIMS = [1024 * 4 ** k for k in range(7)]                      # 1K .. 4M rows
DIA_WIDTHS = list(range(1, 14, 2)) + list(range(17, 62, 4))   # as bench_dia.cpp
FLAT_WIDTHS = [1, 2, 4, 8, 16, 32, 64, 128]                  # bench_ell / sell / coo
CSR_WIDTHS = [2, 8, 32, 128, 512, 2048]                      # bench_csr
BLOCK_COUNTS = list(range(1, 14, 2)) + list(range(17, 30, 4))  # bench_bell / sbell
BCSR_COUNTS = [2, 4, 8, 16, 32, 64, 128, 256, 512]          # bench_bcsr
SHAPES = [(bh, bw) for bw in (4, 8) for bh in (1, 2, 4, 8)]  # analyzer's shape order
OVER_GROUPS = [256 * 2 ** k for k in range(10)]
OVER_THREADS = [64, 128, 256]
CUS = 256

MAX_FLAT = 64 << 20    # stored entries per record
MAX_BLOCK = 24 << 20   # stored block entries per record (built through the conversions)


class _Held:
    """A format struct whose arrays are numpy buffers kept alive here."""

    def __init__(self, s, *arrays):
        self.s = s
        self._keep = arrays


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def band_start(dim, k):
    """First column of each row's k consecutive columns, centred on the
    diagonal and clamped inside the matrix (the tables' row pattern)."""
    return np.clip(np.arange(dim, dtype=np.int64) - k // 2, 0, dim - k).astype(np.int32)


def make_dia(dim, nd, bands=False):
    h = nd // 2
    offs = np.arange(-h, -h + nd, dtype=np.int32)
    r = np.arange(dim, dtype=np.int64)
    data = np.zeros(nd * dim, np.float32)
    nnz = 0
    for d, o in enumerate(offs):
        ok = (r + o >= 0) & (r + o < dim)
        data[d * dim:(d + 1) * dim] = ok
        nnz += int(ok.sum())
    if not bands:
        s = L.Dia()
        s.matinfo.width = s.matinfo.height = dim
        s.matinfo.nnz = nnz
        s.dia_num, s.dia_length, s.dia_length_aligned = nd, dim, dim
        s.dia_offsets, s.dia_data = _ip(offs), _fp(data)
        return _Held(s, offs, data), nnz
    boff = np.array([offs[0]], np.int32)
    bptr = np.array([0, nd], np.int32)
    s = L.Bdia()
    s.matinfo.width = s.matinfo.height = dim
    s.matinfo.nnz = nnz
    s.bdia_band_num, s.bdia_length, s.bdia_length_aligned = 1, dim, dim
    s.bdia_offsets, s.bdia_bptr, s.bdia_data = _ip(boff), _ip(bptr), _fp(data)
    return _Held(s, boff, bptr, data), nnz


def make_ell(dim, k):
    st = band_start(dim, k)
    col = (st[None, :] + np.arange(k, dtype=np.int32)[:, None]).astype(np.int32).ravel()  # slot-major
    val = np.ones(dim * k, np.float32)
    s = L.Ell()
    s.matinfo.width = s.matinfo.height = dim
    s.matinfo.nnz = dim * k
    s.ell_num, s.ell_height_aligned = k, dim
    s.ell_col_id, s.ell_data = _ip(col), _fp(val)
    return _Held(s, col, val), dim * k


def make_sell(dim, k, S):
    st = band_start(dim, k)
    ns = dim // S
    col = np.empty(dim * k, np.int32)
    r = np.arange(dim)
    for slot in range(k):
        col[(r // S) * S * k + slot * S + r % S] = st + slot
    ptr = (np.arange(ns + 1, dtype=np.int64) * S * k).astype(np.int32)
    val = np.ones(dim * k, np.float32)
    s = L.Sell()
    s.matinfo.width = s.matinfo.height = dim
    s.matinfo.nnz = dim * k
    s.sell_slice_height, s.sell_slice_num = S, ns
    s.sell_slice_ptr, s.sell_col_id, s.sell_data = _ip(ptr), _ip(col), _fp(val)
    return _Held(s, ptr, col, val), dim * k


def make_csr(dim, k):
    st = band_start(dim, k)
    col = (st[:, None] + np.arange(k, dtype=np.int32)[None, :]).astype(np.int32).ravel()
    rp = (np.arange(dim + 1, dtype=np.int64) * k).astype(np.int32)
    val = np.ones(dim * k, np.float32)
    s = L.Csr()
    s.matinfo.width = s.matinfo.height = dim
    s.matinfo.nnz = dim * k
    s.csr_row_ptr, s.csr_col_id, s.csr_data = _ip(rp), _ip(col), _fp(val)
    return _Held(s, rp, col, val), dim * k


def make_coo(dim, k):
    st = band_start(dim, k)
    col = (st[:, None] + np.arange(k, dtype=np.int32)[None, :]).astype(np.int32).ravel()
    row = np.repeat(np.arange(dim, dtype=np.int32), k)
    val = np.ones(dim * k, np.float32)
    s = L.Coo()
    s.matinfo.width = s.matinfo.height = dim
    s.matinfo.nnz = dim * k
    s.coo_row_id, s.coo_col_id, s.coo_data = _ip(row), _ip(col), _fp(val)
    return _Held(s, row, col, val), dim * k


def block_coo(dim, bh, bw, nb):
    """Sorted COO of dense bh x bw blocks, nb consecutive block columns per
    block row centred on the diagonal."""
    brows, bcols = dim // bh, dim // bw
    st = np.clip(np.arange(brows, dtype=np.int64) * bh // bw - nb // 2, 0, bcols - nb)
    i = np.arange(brows)[:, None, None, None]
    h = np.arange(bh)[None, :, None, None]
    b = np.arange(nb)[None, None, :, None]
    w = np.arange(bw)[None, None, None, :]
    row = np.broadcast_to(i * bh + h, (brows, bh, nb, bw)).astype(np.int32).ravel()
    col = ((st[:, None, None, None] + b) * bw + w)
    col = np.broadcast_to(col, (brows, bh, nb, bw)).astype(np.int32).ravel()
    return M.CooMatrix.from_arrays(row, col, np.ones(row.size, np.float32), dim, dim)


class Profiler:
    def __init__(self, ctx, quick=False, log=print):
        self.ctx, self.quick, self.log = ctx, quick, log
        n = max(DIMS)
        self.x = M.DeviceArray.from_numpy(ctx, np.ones(n, np.float32))
        self.y = M.DeviceArray.from_numpy(ctx, np.zeros(n, np.float32))

    def gflops(self, fmt, mat, method, nnz):
        part = M.Part(self.ctx, fmt, mat, method)
        try:
            part.spmv(self.x, self.y, 2)
            self.ctx.sync()
            self.ctx.time_begin()
            part.spmv(self.x, self.y, 3)
            ms = self.ctx.time_end() / 3
            n = int(min(200, max(3, (0.4 if self.quick else 2.0) / max(ms, 1e-4))))
            self.ctx.time_begin()
            part.spmv(self.x, self.y, n)
            ms = self.ctx.time_end() / n
        finally:
            part.free()
        return 2.0 * nnz / (ms * 1e-3) / 1e9

    # -------------------------------------------------------------- tables

    def overhead(self):
        rows = []
        us = C.c_float()
        for g in OVER_GROUPS:
            for t in OVER_THREADS:
                check(lib().clspmv_probe_launch_grid(self.ctx.h, g, t, 200, C.byref(us)), "probe_launch_grid")
                rows.append(f"{g} {t} {us.value:f}")
        return rows

    def dia(self, bands):
        rows = []
        for dim in DIMS:
            for nd in DIA_WIDTHS:
                if nd * dim > MAX_FLAT or nd > dim:
                    continue
                mat, nnz = make_dia(dim, nd, bands)
                fmt = L.FMT_BDIA if bands else L.FMT_DIA
                g = [self.gflops(fmt, mat, m, nnz) for m in (0, 1)]
                g += [0.0] * (3 if bands else 2)
                rows.append(f"{dim} {nd} " + " ".join(f"{v:f}" for v in g))
        return rows

    def flat(self, kind):
        rows = []
        widths = CSR_WIDTHS if kind == "csr" else FLAT_WIDTHS
        for dim in DIMS:
            for k in widths:
                if k * dim > MAX_FLAT or k > dim:
                    continue
                if kind == "ell":
                    mat, nnz = make_ell(dim, k)
                    g = [self.gflops(L.FMT_ELL, mat, m, nnz) for m in (0, 1)]
                elif kind == "sell":
                    g = []
                    for S in (32, 128):
                        mat, nnz = make_sell(dim, k, S)
                        g.append(self.gflops(L.FMT_SELL, mat, 0, nnz))
                elif kind == "csr":
                    mat, nnz = make_csr(dim, k)
                    g = [self.gflops(L.FMT_CSR_VECTOR, mat, 1, nnz), 0.0]
                else:
                    mat, nnz = make_coo(dim, k)
                    g = [self.gflops(L.FMT_COO, mat, 0, nnz)]
                lead = f"{dim} {k}" + (f" {CUS}" if kind in ("csr", "coo") else "")
                rows.append(lead + " " + " ".join(f"{v:f}" for v in g))
        return rows

    def block(self, kind):
        rows = []
        counts = BCSR_COUNTS if kind == "bcsr" else BLOCK_COUNTS
        for bh, bw in SHAPES:
            t0 = time.perf_counter()
            for dim in DIMS:
                for nb in counts:
                    if nb * bw > dim or (dim // bh) * nb * bh * bw > MAX_BLOCK:
                        continue
                    coo = block_coo(dim, bh, bw, nb)
                    try:
                        if kind == "bell":
                            mat, fmt = M.coo2b4ell(coo, bw, bh), L.FMT_BELL
                        elif kind == "sbell":
                            mat, fmt = M.coo2sbell(coo, bw, bh, 32), L.FMT_SBELL
                        else:
                            mat, fmt = M.coo2b4csr(coo, bw, bh), L.FMT_BCSR
                    except L.ClspmvError:  # above MAX_MEM_OBJ, as the reference refuses it
                        coo.free()
                        continue
                    g = [self.gflops(fmt, mat, 0, coo.nnz)] + ([0.0] if kind == "bell" else [])
                    mat.free()
                    coo.free()
                    rows.append(f"{bh} {bw} {dim} {nb} " + " ".join(f"{v:f}" for v in g))
            self.log(f"[profiler]   {kind} {bh}x{bw}: {time.perf_counter() - t0:.1f} s")
        return rows

    def all_tables(self):
        out = {}
        plan = [("overhead", self.overhead), ("dia", lambda: self.dia(False)), ("bdia", lambda: self.dia(True)),
                ("ell", lambda: self.flat("ell")), ("sell", lambda: self.flat("sell")),
                ("csr", lambda: self.flat("csr")), ("coo", lambda: self.flat("coo")),
                ("bell", lambda: self.block("bell")), ("sbell", lambda: self.block("sbell")),
                ("bcsr", lambda: self.block("bcsr"))]
        for name, fn in plan:
            t0 = time.perf_counter()
            out[name] = fn()
            self.log(f"[profiler] {name}.ben: {len(out[name])} records in {time.perf_counter() - t0:.1f} s")
        return out


def write_tables(tables, out_dir, note):
    os.makedirs(out_dir, exist_ok=True)
    for name, rows in tables.items():
        with open(os.path.join(out_dir, f"{name}.ben"), "w") as f:
            f.write("\n".join(rows) + "\n")
    with open(os.path.join(out_dir, "README.md"), "w") as f:
        f.write(note)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.abspath(__file__)), "tables", "mi355x"))
    ap.add_argument("--quick", action="store_true", help="shorter timing loops")
    args = ap.parse_args(argv)
    ctx = M.Context(0)
    t0 = time.perf_counter()
    prof = Profiler(ctx, args.quick, log=lambda s: print(s, file=sys.stderr, flush=True))
    tables = prof.all_tables()
    note = (f"# MI355X tables\n\nMeasured by `python -m clspmv_amd.profiler` on one MI355X (gfx950), "
            f"{datetime.date.today().isoformat()}, {time.perf_counter() - t0:.0f} s.  Same schema as the "
            f"reference's tables; see the module docstring of clspmv_amd/profiler.py for the matrices and "
            f"the meaning of every column.\n")
    write_tables(tables, args.out, note)
    ctx.close()
    print(f"wrote {sum(len(v) for v in tables.values())} records to {args.out}", flush=True)


if __name__ == "__main__":
    main()
