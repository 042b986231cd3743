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
    for name in args.matrices.split(","):
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


if __name__ == "__main__":
    main()
