"""Kernel profiling driver: runs selected SpMV kernels a few times each plus a
plain device copy of known size, so that rocprofv3 counter runs
(FETCH_SIZE, WRITE_SIZE, TCC_HIT/MISS ...) can be scaled by the copy's
readings.

    python -m clspmv_amd.kprof --matrices c1a,c1b,c1c,c1d --ntimes 5 --copy-mib 1024
Prints one JSON line with each kernel's name, launches and algorithmic bytes.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import sys

import numpy as np

from . import _lib as L
from . import matrix as M
from . import workloads as W


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--matrices", default="c1a,c1b,c1c,c1d")
    ap.add_argument("--ntimes", type=int, default=5)
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--methods", default="", help="kernel variant overrides, e.g. c1b=2,c1c=2")
    ap.add_argument("--scale", type=int, default=1)
    ap.add_argument("--batch-blocks", type=int, default=0, metavar="N",
                    help="also run one batch of N 3x3-block matrices (BSR entries, k = 16: the matrix-core "
                         "kernel) ntimes + 1 times")
    args = ap.parse_args(argv)
    ctx = M.Context(0)
    out = {"kernels": [], "copy": None}
    # plain copy of known size: reads and writes copy_mib MiB once per launch
    nbytes = args.copy_mib << 20
    src = M.DeviceArray(ctx, nbytes // 4)
    dst = M.DeviceArray(ctx, nbytes // 4)
    g = C.c_float()
    L.check(L.lib().clspmv_probe_copy(ctx.h, src.ptr, dst.ptr, nbytes, args.ntimes, C.byref(g)), "probe_copy")
    out["copy"] = {"kernel": "copy_probe_kernel", "bytes_read": nbytes, "bytes_written": nbytes,
                   "launches": args.ntimes + 1, "gbps": round(g.value, 1)}
    src.free()
    dst.free()
    overrides = {}
    for item in filter(None, args.methods.split(",")):
        k, v = item.split("=")
        overrides.setdefault(k, []).append(int(v))
    if args.batch_blocks:
        out["batch"] = _run_batch(ctx, args.batch_blocks, args.ntimes)
    for name in filter(None, args.matrices.split(",")):
        wl = W.build(name, args.scale)
        for meth in overrides.get(name, [wl.method]):
            _run_one(ctx, out, name, wl, meth, args.ntimes)
        wl.mat.free()
    ctx.close()
    print(json.dumps(out), flush=True)


def _run_one(ctx, out, name, wl, meth, ntimes):
    part = M.Part(ctx, wl.fmt, wl.mat, meth)
    dx = M.DeviceArray.from_numpy(ctx, M.gen_vector(wl.coo.width, 7))
    dy = M.DeviceArray.from_numpy(ctx, np.zeros(wl.coo.height, np.float32))
    part.spmv(dx, dy, 1)
    ctx.sync()
    ctx.time_begin()
    part.spmv(dx, dy, ntimes)
    ms = ctx.time_end() / ntimes
    out["kernels"].append({"matrix": name, "kernel": part.info["kernel"], "method": meth,
                           "launches_per_spmv": part.info["launches"], "spmvs": ntimes + 1,
                           "alg_bytes": part.info["alg_bytes"], "stored_bytes": part.info["stored_bytes"],
                           "nnz": part.info["nnz"], "us": round(ms * 1e3, 2),
                           "gbps_alg": round(part.info["alg_bytes"] / (ms * 1e-3) / 1e9, 1)})
    print(f"[kprof] {name} m{meth} {part.info['kernel']} {ms * 1e3:.1f} us "
          f"{part.info['alg_bytes'] / (ms * 1e-3) / 1e9:.0f} GB/s", file=sys.stderr, flush=True)
    part.free()
    dx.free()
    dy.free()


def _run_batch(ctx, n, ntimes, k=16):
    """One batch of n matrices of 3x3 blocks (9 per block row, 65536 block
    rows each) as BSR entries with k = 16 right-hand sides: the batched
    kernel's block items on the matrix cores."""
    from . import batch as B
    coos = [M.gen_block_rows(65536, 3, 9, 5000, 900 + i) for i in range(n)]
    mats = [M.coo2bsr(c, 3, 3) for c in coos]
    b = B.Batch(ctx, mats, [k] * n)
    for i, c in enumerate(coos):
        b.set_x(i, np.tile(M.gen_vector(c.width, 30 + i)[:, None], (1, k)))
    b.run(1)
    ctx.sync()
    ctx.time_begin()
    b.run(ntimes)
    ms = ctx.time_end() / ntimes
    st = b.stats()
    flops = 2.0 * st["nnz_rhs"]
    res = {"matrices": n, "k": k, "nnz": st["nnz"], "nnz_rhs": st["nnz_rhs"], "alg_bytes": st["alg_bytes"],
           "runs": ntimes + 1, "us": round(ms * 1e3, 2), "tflops": round(flops / (ms * 1e-3) / 1e12, 3),
           "gbps_alg": round(st["alg_bytes"] / (ms * 1e-3) / 1e9, 1)}
    print(f"[kprof] batch {n} x 3x3 blocks k={k}: {ms * 1e3:.1f} us, {res['tflops']} TFLOP/s, "
          f"{res['gbps_alg']} GB/s", file=sys.stderr, flush=True)
    b.free()
    for m in mats:
        m.free()
    for c in coos:
        c.free()
    return res


if __name__ == "__main__":
    main()
