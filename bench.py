#!/usr/bin/env python
"""Benchmark of the SpMV engine on one MI355X.

Default workload (c1, the single-format FP32 kernels of the headline metric):
one step = one y += A x on each of four seeded matrices, each in the format it
is meant to exercise, all resident in HBM:
    c1a 27-point stencil 128^3            DIA
    c1b 4M rows x 32 uniform columns      ELL, LDS x-tiled layout
    c1c 2M rows Zipf(1.8) mean 16         CSR, LDS x-tiled layout
    c1d 1M block rows of 3x3 blocks x 9   BSR 3x3, sliced layout
value = sum(2*nnz) / step time (GFLOP/s).  Each kernel is also timed alone
with HIP events on the context stream; the dominant one is reported against
the HBM roofline with its algorithmic bytes (format arrays + x + y read and
write).  The Cocktail composite (c2, 4M rows) is then analyzed, uploaded and
timed beside the best single format (extra field "cocktail_c2"), the c1
matrices are run again with mixed-precision values (field "mixed_c3": time,
bytes, speedup and error against an FP64 product per codec), and a batch of
matrices with 1-16 right-hand sides goes through one captured graph (field
"batched_c4").  The
reference's CPU SpMV (spmv_only -> coo_spmv, one core) is timed on the host
on a bounded sample of the same c1 matrices.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-c2] [--no-c3] [--no-cpu]
                       [--c4-matrices N]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
METRIC = "SpMV GFLOP/s + HBM GB/s (% of peak), mixed/batched vs plain FP32 Cocktail"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(workloads, budget_s=12.0):
    """The reference's CPU SpMV (spmv_only -> coo_spmv, A/src/util.cpp:85-91,
    A/include/spmv_serial.h:184-196) as restated in oracle/spmv_oracle.c
    (bit-exact with the reference's compiled spmv_only on the golden
    matrices), one core, on a bounded sample of the same matrices."""
    from oracle import oracle as O

    total_flops = 0.0
    total_t = 0.0
    reps_done = []
    for w in workloads:
        coo = w["coo"]
        x = w["x"]
        reps = 0
        t_w = 0.0
        while reps == 0 or (t_w < budget_s / len(workloads) and reps < 100):
            t0 = time.perf_counter()
            O.spmv_only(coo.row_id, coo.col_id, coo.data, x, coo.height)
            t_w += time.perf_counter() - t0
            reps += 1
        total_t += t_w
        total_flops += 2.0 * coo.nnz * reps
        reps_done.append(f"{w['name']}x{reps}")
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    out = {
        "value": round(total_flops / total_t / 1e9, 4),
        "unit": "GFLOP/s",
        "cores": 1,
        "kind": "port",
        "sample": "oracle_spmv_only, the restated reference spmv_only (y=0; y[row]+=val*x[col] in COO order, "
                  "FP32, no FMA), one thread, over the full matrices: "
                  + ", ".join(reps_done) + f"; {total_t:.1f} s; host {os.cpu_count()} CPUs, {model}",
    }
    # project CPU baseline (not the reference): OpenMP CSR on all host cores
    from clspmv_amd import matrix as M

    flops, t_all, threads, done = 0.0, 0.0, 1, []
    for w in workloads:
        csr = M.coo2csr(w["coo"])
        reps, t_w = 0, 0.0
        while reps == 0 or (t_w < budget_s / 4 / len(workloads) and reps < 100):
            t0 = time.perf_counter()
            _, threads = O.csr_spmv_omp(csr.row_ptr, csr.col_id, csr.data, w["x"])
            t_w += time.perf_counter() - t0
            reps += 1
        csr.free()
        t_all += t_w
        flops += 2.0 * w["coo"].nnz * reps
        done.append(f"{w['name']}x{reps}")
    out["project_omp"] = {"value": round(flops / t_all / 1e9, 3), "unit": "GFLOP/s", "cores": threads,
                          "kind": "project CPU, not reference",
                          "sample": "OpenMP CSR rows, FP32, over the full matrices: " + ", ".join(done)
                                    + f"; {t_all:.1f} s"}
    return out


def cocktail_c2(ctx, args):
    """The Cocktail composite (4M rows: band + 4x4 blocks + Zipf rows):
    analyze with the reference's performance tables and with the MI355X
    tables, upload every part, time do_spmv; beside it the best single
    format (CSR, x-tiled or column-sliced) on the same matrix."""
    import numpy as np

    from clspmv_amd import _lib as L
    from clspmv_amd import cocktail as CK
    from clspmv_amd import matrix as M
    from clspmv_amd import workloads as W

    coo = W.make_coo("c2", args.scale)
    x = M.gen_vector(coo.width, 7)
    k = max(10, args.steps)
    out = {"rows": coo.height, "nnz": int(coo.nnz)}
    for label, path in (("reference_tables", CK.REFERENCE_TABLES), ("mi355x_tables", CK.MI355X_TABLES)):
        t0 = time.perf_counter()
        mat = CK.analyze_matrix(coo.copy(), CK.BenchTables(path))
        t_an = time.perf_counter() - t0
        gpu = CK.GpuCocktail(ctx, mat, x, np.zeros(coo.height, np.float32))
        parts = gpu.part_infos()
        gpu.do_spmv(2)
        ctx.sync()
        ctx.time_begin()
        gpu.do_spmv(k)
        ms = ctx.time_end() / k
        alg = sum(p["stored_bytes"] for p in parts) + 4 * coo.width + 8 * coo.height
        out[label] = {
            "analysis_s": round(t_an, 2), "partition": mat.describe()["parts"],
            "kernels": [p["kernel"] for p in parts],
            "ms": round(ms, 5), "gflops": round(2.0 * coo.nnz / (ms * 1e-3) / 1e9, 1),
            "alg_bytes": int(alg), "gbps": round(alg / (ms * 1e-3) / 1e9, 1),
            "pct_hbm_peak": round(100.0 * alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 1),
        }
        log(f"c2 Cocktail ({label}): {out[label]['partition']} {out[label]['kernels']} {ms * 1e3:.1f} us "
            f"{out[label]['gflops']} GFLOP/s, analysis {t_an:.1f} s")
        gpu.free()
        mat.free()
    best = None
    csr = M.coo2csr(coo)
    dx = M.DeviceArray.from_numpy(ctx, x)
    dy = M.DeviceArray.from_numpy(ctx, np.zeros(coo.height, np.float32))
    for meth in (4, 2):
        part = M.Part(ctx, L.FMT_CSR_VECTOR, csr, meth)
        part.spmv(dx, dy, 2)
        ctx.sync()
        ctx.time_begin()
        part.spmv(dx, dy, k)
        m = ctx.time_end() / k
        if best is None or m < best["ms"]:
            best = {"format": "CSR", "kernel": part.info["kernel"], "ms": round(m, 5),
                    "gflops": round(2.0 * coo.nnz / (m * 1e-3) / 1e9, 1)}
        part.free()
    csr.free()
    dx.free()
    dy.free()
    out["best_single_format"] = best
    for label in ("reference_tables", "mi355x_tables"):
        out[label]["vs_best_single"] = round(best["ms"] / out[label]["ms"], 3)
    log(f"c2 best single format: {best['kernel']} {best['ms'] * 1e3:.1f} us")
    return out


def _fp64_reference(coo, x):
    """FP64 y = A x and the row sums S = |A| |x| (scipy, on the host)."""
    import numpy as np
    import scipy.sparse as sp

    a = sp.csr_matrix((coo.data.astype(np.float64), (coo.row_id, coo.col_id)), shape=(coo.height, coo.width))
    x64 = x.astype(np.float64)
    y64 = a @ x64
    a.data = np.abs(a.data)
    return y64, a @ np.abs(x64)


def _errors(y, y64, s):
    """max_i |y_i - y64_i| / max(|y64_i|, S_i) (componentwise, S = |A||x|) and
    ||y - y64||_2 / ||y64||_2."""
    import numpy as np

    d = np.abs(y.astype(np.float64) - y64)
    return (float(np.max(d / np.maximum(np.maximum(np.abs(y64), s), 1e-300))),
            float(np.linalg.norm(d) / max(np.linalg.norm(y64), 1e-300)))


def mixed_c3(ctx, args, work, per):
    """Mixed precision (c3) on the c1 matrices and the c2 composite: each c1
    matrix in its c1 layout with values stored as BF16 hi+lo, FP16 with a
    row scale or FP8 E4M3 with a slice scale (DIA and x-tiled CSR parts; BSR
    with int16 delta block columns), FP32 accumulation; c2 through its
    Cocktail partition with the codec on the parts that have one (x-tiled
    and sliced BSR row runs) against the FP32 partition.  Per matrix and codec:
    kernel time (HIP events), stored bytes, effective nonzeros/s, speedup
    over FP32, and the error of one y = A x against the FP64 product."""
    import numpy as np

    from clspmv_amd import _lib as L
    from clspmv_amd import matrix as M

    k = max(10, args.steps)
    out = {}

    def run_once(part, dx, n):
        dy = M.DeviceArray.from_numpy(ctx, np.zeros(n, np.float32))
        part.spmv(dx, dy, 1)
        y = dy.download()
        dy.free()
        return y

    def time_part(part, dx, dy, copies=()):
        """HIP-event time per launch, rotating over (part, x, y) copies when
        given (a working set above the Infinity Cache)."""
        rot = [(part, dx, dy)] + list(copies)
        for p_, x_, y_ in rot:
            p_.spmv(x_, y_, 1)
        ctx.sync()
        kr = max(1, k // len(rot)) * len(rot)
        ctx.time_begin()
        for i in range(kr):
            p_, x_, y_ = rot[i % len(rot)]
            p_.spmv(x_, y_, 1)
        return ctx.time_end() / kr

    for w in work:
        coo, name = w["coo"], w["name"]
        y64, srow = _fp64_reference(coo, w["x"])
        e_max, e_norm = _errors(run_once(w["part"], w["dx"], coo.height), y64, srow)
        ms32 = per[name]["ms"]
        rec = {"fp32": {"kernel": w["part"].info["kernel"], "ms": ms32,
                        "stored_bytes": int(w["part"].info["stored_bytes"]),
                        "eff_gnnz_per_s": round(coo.nnz / (ms32 * 1e-3) / 1e9, 1),
                        "max_rel_err": e_max, "norm_rel_err": e_norm}}
        if name == "c1a":
            fmt, mat = L.FMT_DIA, M.coo2dia(coo, 256)
        elif name == "c1d":
            fmt, mat = L.FMT_BSR, M.coo2bsr(coo, 3, 3)
        else:
            fmt, mat = L.FMT_CSR_VECTOR, M.coo2csr(coo)
        _mixed_codecs(ctx, M, L, rec, fmt, mat, coo, w, ms32, y64, srow, time_part, run_once)
        mat.free()
        out[name] = rec
    if args.no_c2:
        return out
    # c2: the Cocktail composite; FP32 = the Cocktail object with the MI355X
    # tables; mixed = the same partition with the value codec on its parts
    # that have a mixed-precision variant (x-tiled and sliced BSR row runs)
    import ctypes as C

    from clspmv_amd import cocktail as CK
    from clspmv_amd import workloads as W

    coo = W.make_coo("c2", args.scale)
    x = M.gen_vector(coo.width, 7)
    y64, srow = _fp64_reference(coo, x)
    mat = CK.analyze_matrix(coo.copy(), CK.BenchTables(CK.MI355X_TABLES))
    rec = {}
    ms32 = None
    for codec, label in ((0, "fp32"),) + _codec_labels(L):
        L.check(L.lib().clspmv_set_option(b"cocktail_value_codec", C.c_longlong(codec)), "set_option")
        try:
            gpu = CK.GpuCocktail(ctx, mat, x, np.zeros(coo.height, np.float32))
        finally:
            L.check(L.lib().clspmv_set_option(b"cocktail_value_codec", C.c_longlong(0)), "set_option")
        gpu.do_spmv(2)
        ctx.sync()
        ctx.time_begin()
        gpu.do_spmv(k)
        ms = ctx.time_end() / k
        gpu.result_from_cpu(np.zeros(coo.height, np.float32))
        gpu.do_spmv(1)
        e_max, e_norm = _errors(gpu.result_to_cpu(), y64, srow)
        infos = gpu.part_infos()
        ms32 = ms if codec == 0 else ms32
        # the Cocktail driver's error guard (default 1e-4) keeps a run's
        # codec only when its codec_bound allows; the others run in FP32
        rec[label] = {"kernel": "Cocktail: " + "+".join(p["kernel"] for p in infos), "ms": round(ms, 5),
                      "stored_bytes": int(sum(p["stored_bytes"] for p in infos)),
                      "eff_gnnz_per_s": round(coo.nnz / (ms * 1e-3) / 1e9, 1),
                      "max_rel_err": e_max, "norm_rel_err": e_norm,
                      "codec_bound": max([p["codec_bound"] for p in infos] + [0.0]),
                      "parts_with_codec": sum(1 for p in infos if p["codec_bound"] > 0 or "mixed" in p["kernel"])}
        if codec:
            rec[label]["speedup_vs_fp32"] = round(ms32 / ms, 3)
            log(f"c3 c2 {rec[label]['kernel']}: {ms * 1e3:.1f} us ({ms32 / ms:.2f}x FP32 Cocktail), "
                f"max rel err {e_max:.2e}, norm rel err {e_norm:.2e}")
        gpu.free()
    _best_within(rec, 1e-4)
    mat.free()
    coo.free()
    out["c2"] = rec
    return out


def _codec_labels(L):
    return ((L.VAL_BF16X2, "bf16x2"), (L.VAL_FP16_ROW, "fp16row"), (L.VAL_FP8_SLICE, "fp8slice"),
            (L.VAL_I16_ROW, "i16row"))


def _mixed_codecs(ctx, M, L, rec, fmt, mat, coo, w, ms32, y64, srow, time_part, run_once, name=None):
    """Adds one record per value codec to rec (time, bytes, speedup, errors)."""
    import numpy as np

    for codec, label in _codec_labels(L):
        part = M.MixedPart(ctx, fmt, mat, codec, True)
        # copies so that the timed launches rotate over > 600 MB
        work_set = int(part.info["stored_bytes"]) + 4 * coo.width + 8 * coo.height
        copies = [(M.MixedPart(ctx, fmt, mat, codec, True), M.DeviceArray.from_numpy(ctx, w["x"]),
                   M.DeviceArray.from_numpy(ctx, np.zeros(coo.height, np.float32)))
                  for _ in range(max(0, min(4, -(-600_000_000 // work_set) - 1)))]
        ms = time_part(part, w["dx"], w["dy"], copies)
        for p_, x_, y_ in copies:
            p_.free()
            x_.free()
            y_.free()
        e_max, e_norm = _errors(run_once(part, w["dx"], coo.height), y64, srow)
        rec[label] = {"kernel": part.info["kernel"], "ms": round(ms, 5),
                      "stored_bytes": int(part.info["stored_bytes"]),
                      "eff_gnnz_per_s": round(coo.nnz / (ms * 1e-3) / 1e9, 1),
                      "speedup_vs_fp32": round(ms32 / ms, 3), "max_rel_err": e_max, "norm_rel_err": e_norm,
                      "codec_bound": part.info["codec_bound"]}
        log(f"c3 {name or w.get('name')} {part.info['kernel']}: {ms * 1e3:.1f} us ({ms32 / ms:.2f}x FP32), "
            f"{part.info['stored_bytes'] / 1e6:.1f} MB, max rel err {e_max:.2e}, norm rel err {e_norm:.2e}")
        part.free()
    _best_within(rec, 1e-4)


def _best_within(rec, tol):
    """rec["best_max_err_1e-4"]: the fastest codec whose measured
    componentwise error (this x) is at most tol; rec["best_guaranteed_1e-4"]:
    the fastest whose codec_bound -- the encoding's bound on the
    componentwise error for any x -- is at most tol (each with its speedup
    over FP32; absent when no codec qualifies)."""
    for key, ok in (("best_max_err_1e-4", lambda r: r["max_rel_err"] <= tol),
                    ("best_guaranteed_1e-4", lambda r: r.get("codec_bound", 1.0) <= tol and r["max_rel_err"] <= tol)):
        cand = [(r["ms"], lab) for lab, r in rec.items() if lab != "fp32" and isinstance(r, dict) and "ms" in r
                and r.get("parts_with_codec", 1) > 0 and ok(r)]
        if cand:
            ms, lab = min(cand)
            rec[key] = {"codec": lab, "speedup_vs_fp32": rec[lab]["speedup_vs_fp32"]}


def batched_c4(ctx, args):
    """The batch workload (c4): n matrices, k in {1, 4, 16}.  Three ways to
    compute every Y_i += A_i X_i, timed with HIP events, in nonzeros x
    right-hand sides per second:
      plain     k_i single-vector FP32 SpMVs per matrix, CSR adaptive row
                blocks (one launch each);
      selector  the same with the format the learned selector picks;
      batched   one replay of the captured graph (persistent kernel over all
                matrices);
    and, on a seeded eighth of the matrices, the plain FP32 Cocktail path
    (analyzer + k do_spmv per matrix) against a batch of the same
    matrices."""
    import numpy as np

    from clspmv_amd import batch as B
    from clspmv_amd import _lib as L
    from clspmv_amd import matrix as M
    from clspmv_amd import selector as S
    from clspmv_amd import workloads as W

    t0 = time.perf_counter()
    mats = W.c4_matrices(args.c4_matrices)
    csrs = [M.coo2csr(coo) for _, coo, _ in mats]
    ks = [k for _, _, k in mats]
    nnz = sum(coo.nnz for _, coo, _ in mats)
    nnz_rhs = sum(coo.nnz * k for _, coo, k in mats)
    log(f"c4: {len(mats)} matrices, {nnz} nnz, {nnz_rhs} nnz x rhs ({time.perf_counter() - t0:.1f}s setup)")
    out = {"matrices": len(mats), "nnz": int(nnz), "nnz_x_rhs": int(nnz_rhs),
           "rows_log2_range": [10, 20], "k_cycle": [1, 4, 16]}
    reps = 3

    # batched: one graph launch
    b = B.Batch(ctx, csrs, ks)
    for i, (_, coo, k) in enumerate(mats):
        b.set_x(i, np.ones((coo.width, k), np.float32))
    b.capture()
    b.run(1)
    ctx.sync()
    ctx.time_begin()
    b.run(reps)
    ms_b = ctx.time_end() / reps
    st = b.stats()
    out["batched"] = {"ms": round(ms_b, 4), "eff_gnnz_per_s": round(nnz_rhs / (ms_b * 1e-3) / 1e9, 1),
                      "alg_gbps": round(st["alg_bytes"] / (ms_b * 1e-3) / 1e9, 1), "items": st["items"],
                      "launch": "one hipGraph replay"}
    b.free()

    def per_matrix(choose, only=None):
        parts = []
        for i, ((kind, coo, k), csr) in enumerate(zip(mats, csrs)):
            if only is not None and i not in only:
                continue
            fmt, meth, mat = choose(coo, csr)
            part = M.Part(ctx, fmt, mat, meth)
            if mat is not csr and mat is not coo:
                mat.free()
            x = M.DeviceArray.from_numpy(ctx, np.ones(coo.width, np.float32))
            y = M.DeviceArray(ctx, coo.height)
            parts.append((part, x, y, k))
        for part, x, y, k in parts:
            part.spmv(x, y, k)
        ctx.sync()
        ctx.time_begin()
        for _ in range(reps):
            for part, x, y, k in parts:
                part.spmv(x, y, k)
        ms = ctx.time_end() / reps
        for part, x, y, _ in parts:
            part.free()
            x.free()
            y.free()
        return ms

    # the same split by k (each class alone), so the source of the gain is visible
    by_k = {}
    for kk in (1, 4, 16):
        idx = [i for i, k in enumerate(ks) if k == kk]
        if not idx:
            continue
        sub = B.Batch(ctx, [csrs[i] for i in idx], [kk] * len(idx))
        for j, i in enumerate(idx):
            sub.set_x(j, np.ones((mats[i][1].width, kk), np.float32))
        sub.capture()
        sub.run(1)
        ctx.sync()
        ctx.time_begin()
        sub.run(reps)
        ms_bk = ctx.time_end() / reps
        sub.free()
        nz = sum(mats[i][1].nnz for i in idx)
        by_k[str(kk)] = {"matrices": len(idx), "nnz_x_rhs": int(nz * kk), "batched_ms": round(ms_bk, 4),
                         "batched_eff_gnnz_per_s": round(nz * kk / (ms_bk * 1e-3) / 1e9, 1)}

    ms_p = per_matrix(lambda coo, csr: (L.FMT_CSR_VECTOR, 1, csr))
    out["plain"] = {"ms": round(ms_p, 4), "eff_gnnz_per_s": round(nnz_rhs / (ms_p * 1e-3) / 1e9, 1),
                    "kernel": "csr_adaptive, k launches per matrix"}
    for kk, rec in by_k.items():
        idx = {i for i, k in enumerate(ks) if k == int(kk)}
        ms_pk = per_matrix(lambda coo, csr: (L.FMT_CSR_VECTOR, 1, csr), idx)
        rec["plain_ms"] = round(ms_pk, 4)
        rec["plain_eff_gnnz_per_s"] = round(rec["nnz_x_rhs"] / (ms_pk * 1e-3) / 1e9, 1)
        rec["batched_vs_plain"] = round(ms_pk / rec["batched_ms"], 3)
        log(f"c4 k={kk}: {rec['matrices']} matrices, plain {ms_pk:.2f} ms, batched {rec['batched_ms']:.2f} ms "
            f"({rec['batched_vs_plain']}x)")
    out["by_k"] = by_k
    sel = S.Selector.load_default(ctx)
    choice = {}

    def pick(coo, csr):
        fmt, meth, name = sel.choose(coo)
        choice[id(coo)] = name
        key = S.CANDIDATES[S.CANDIDATE_NAMES.index(name)][3]
        try:
            mat = S._convert(key, coo) if key != "csr" else csr
        except Exception:
            mat = None
        if mat is None:
            return L.FMT_CSR_VECTOR, 1, csr
        return fmt, meth, mat

    ms_s = per_matrix(pick)
    sel.net.free()
    out["selector"] = {"ms": round(ms_s, 4), "eff_gnnz_per_s": round(nnz_rhs / (ms_s * 1e-3) / 1e9, 1),
                       "kernel": "learned selector's format, k launches per matrix"}
    out["batched_vs_plain"] = round(ms_p / ms_b, 3)
    out["selector_vs_plain"] = round(ms_p / ms_s, 3)

    # the combined form: one graph over all matrices, each matrix in the
    # path the learned selector picks -- BSR b x b blocks on the matrix
    # cores where it picks a sliced BSR class and k >= 4, CSR items otherwise
    # (value codecs: none passes the 1e-4 guard on these N(0,1) values)
    def combined_entries(idx):
        mats_c, held = [], []
        for i in idx:
            _, coo, k = mats[i]
            name = choice.get(id(coo), "csr_adaptive")
            if name.startswith("bsr") and k >= 4:
                bd = int(name[3])
                m = M.coo2bsr(coo, bd, bd)
                mats_c.append(m)
                held.append(m)
            else:
                mats_c.append(csrs[i])
        return mats_c, held

    def time_batch(bt, reps_=reps):
        bt.capture()
        bt.run(1)
        ctx.sync()
        ctx.time_begin()
        bt.run(reps_)
        return ctx.time_end() / reps_

    all_idx = list(range(len(mats)))
    mats_c, held = combined_entries(all_idx)
    bc = B.Batch(ctx, mats_c, ks)
    for i, (_, coo, k) in enumerate(mats):
        bc.set_x(i, np.ones((coo.width, k), np.float32))
    ms_cb = time_batch(bc)
    st_c = bc.stats()
    bc.free()
    n_block = sum(1 for i in all_idx if choice.get(id(mats[i][1]), "").startswith("bsr") and ks[i] >= 4)
    out["combined"] = {"ms": round(ms_cb, 4), "eff_gnnz_per_s": round(nnz_rhs / (ms_cb * 1e-3) / 1e9, 1),
                       "alg_gbps": round(st_c["alg_bytes"] / (ms_cb * 1e-3) / 1e9, 1),
                       "matrices_on_matrix_cores": n_block, "vs_plain": round(ms_p / ms_cb, 3),
                       "vs_batched_csr": round(ms_b / ms_cb, 3),
                       "launch": "one hipGraph replay; per matrix the selector's path (BSR MFMA or CSR items)"}
    for kk in (1, 4, 16):
        idx = [i for i in all_idx if ks[i] == kk]
        if not idx:
            continue
        mk, hk = combined_entries(idx)
        bk = B.Batch(ctx, mk, [kk] * len(idx))
        for j, i in enumerate(idx):
            bk.set_x(j, np.ones((mats[i][1].width, kk), np.float32))
        ms_k = time_batch(bk)
        bk.free()
        for m in hk:
            m.free()
        nz = sum(mats[i][1].nnz for i in idx)
        by_k[str(kk)]["combined_ms"] = round(ms_k, 4)
        by_k[str(kk)]["combined_eff_gnnz_per_s"] = round(nz * kk / (ms_k * 1e-3) / 1e9, 1)
        by_k[str(kk)]["combined_vs_plain"] = round(by_k[str(kk)]["plain_ms"] / ms_k, 3)
    log(f"c4 combined (selector's path per matrix, {n_block} on the matrix cores): {ms_cb:.2f} ms "
        f"({out['combined']['vs_plain']}x plain, {out['combined']['vs_batched_csr']}x the CSR batch)")

    # plain FP32 Cocktail per matrix (analyzer with the MI355X tables, k
    # do_spmv per matrix) on an eighth of the matrices, against one graph
    # replay of a batch of the same matrices
    from clspmv_amd import cocktail as CK

    # a seeded eighth of the set (every kind and every k)
    sub = sorted(np.random.default_rng(8).choice(len(mats), max(1, len(mats) // 8), replace=False).tolist())
    tables = CK.BenchTables(CK.MI355X_TABLES)
    t0 = time.perf_counter()
    objs = []
    for i in sub:
        _, coo, k = mats[i]
        cm = CK.analyze_matrix(coo.copy(), tables)
        objs.append((CK.GpuCocktail(ctx, cm, np.ones(coo.width, np.float32), np.zeros(coo.height, np.float32)),
                     cm, k))
    t_an = time.perf_counter() - t0
    for gpu, _, k in objs:
        gpu.do_spmv(k)
    ctx.sync()
    ctx.time_begin()
    for _ in range(reps):
        for gpu, _, k in objs:
            gpu.do_spmv(k)
    ms_c = ctx.time_end() / reps
    for gpu, cm, _ in objs:
        gpu.free()
        cm.free()
    bs = B.Batch(ctx, [csrs[i] for i in sub], [ks[i] for i in sub])
    for j, i in enumerate(sub):
        bs.set_x(j, np.ones((mats[i][1].width, ks[i]), np.float32))
    bs.capture()
    bs.run(1)
    ctx.sync()
    ctx.time_begin()
    bs.run(reps)
    ms_bs = ctx.time_end() / reps
    bs.free()
    # the combined form on the same matrices, and its error: every right-hand
    # side of a seeded sixteenth of them against the FP64 product
    mats_s, held_s = combined_entries(sub)
    bcs = B.Batch(ctx, mats_s, [ks[i] for i in sub])
    rng = np.random.default_rng(9)
    xs = {}
    for j, i in enumerate(sub):
        xs[j] = rng.uniform(-1, 1, (mats[i][1].width, ks[i])).astype(np.float32)
        bcs.set_x(j, xs[j])
    ms_sub_c = time_batch(bcs)
    for j in range(len(sub)):
        bcs.set_y(j, np.zeros((mats[sub[j]][1].height, ks[sub[j]]), np.float32))
    bcs.run(1)
    ctx.sync()
    worst = 0.0
    checked = 0
    for j in range(0, len(sub), 8):
        coo = mats[sub[j]][1]
        y = bcs.get_y(j)
        for r in range(ks[sub[j]]):
            y64, srow = _fp64_reference(coo, xs[j][:, r])
            worst = max(worst, _errors(y[:, r], y64, srow)[0])
            checked += 1
    bcs.free()
    for m in held_s:
        m.free()
    nz_sub = sum(mats[i][1].nnz * ks[i] for i in sub)
    out["cocktail_subset"] = {
        "matrices": len(sub), "nnz_x_rhs": int(nz_sub), "analysis_s": round(t_an, 1),
        "cocktail_ms": round(ms_c, 4), "cocktail_eff_gnnz_per_s": round(nz_sub / (ms_c * 1e-3) / 1e9, 1),
        "batched_ms": round(ms_bs, 4), "batched_eff_gnnz_per_s": round(nz_sub / (ms_bs * 1e-3) / 1e9, 1),
        "batched_vs_cocktail": round(ms_c / ms_bs, 3),
        "combined_ms": round(ms_sub_c, 4), "combined_eff_gnnz_per_s": round(nz_sub / (ms_sub_c * 1e-3) / 1e9, 1),
        "combined_vs_cocktail": round(ms_c / ms_sub_c, 3),
        "combined_max_componentwise_err": worst, "combined_err_vectors_checked": checked,
        "kernel": "Cocktail partition per matrix (MI355X tables), k do_spmv each"}
    log(f"c4 Cocktail per matrix ({len(sub)} matrices, analysis {t_an:.1f} s): {ms_c:.2f} ms, batched "
        f"{ms_bs:.2f} ms ({out['cocktail_subset']['batched_vs_cocktail']}x)")
    for m in held:
        m.free()
    for csr in csrs:
        csr.free()
    for _, coo, _ in mats:
        coo.free()
    log(f"c4: plain {ms_p:.2f} ms, selector {ms_s:.2f} ms, batched {ms_b:.2f} ms "
        f"(batched {out['batched']['eff_gnnz_per_s']} G nnz*rhs/s, {out['batched_vs_plain']}x plain)")
    return out


def launch_c0(ctx):
    """Launch overhead and the hot loop on a small matrix (the 256 x 256
    5-point Laplacian, Cocktail partition with the reference's tables):
    the launch-overhead table (plain launch, captured graph of n empty
    kernels), do_spmv from captured graphs against direct launches, and the
    iteration {y += A x; x := y} on the device against the reference's
    host-driven loop (do_spmv + cpy_vector_from_gpu per step)."""
    import ctypes as C

    import numpy as np

    from clspmv_amd import _lib as L
    from clspmv_amd import cocktail as CK
    from clspmv_amd import matrix as M
    from clspmv_amd import workloads as W

    out = {}
    us = C.c_float()
    L.check(L.lib().clspmv_probe_launch(ctx.h, 2000, C.byref(us)), "probe_launch")
    out["plain_launch_us"] = round(us.value, 3)
    for n in (1, 8, 64):
        L.check(L.lib().clspmv_probe_graph_launch(ctx.h, n, 500, C.byref(us)), "probe_graph_launch")
        out[f"graph_{n}_nodes_us_per_launch"] = round(us.value, 3)
        out[f"graph_{n}_nodes_us_per_node"] = round(us.value / n, 3)
    c0 = W.make_coo("c0")
    # values scaled by 1/20 so that 100 steps of y <- (I + A) y stay finite
    coo = M.CooMatrix.from_arrays(c0.row_id.copy(), c0.col_id.copy(), c0.data * np.float32(0.05), c0.width,
                                  c0.height)
    c0.free()
    mat = CK.analyze_matrix(coo.copy(), CK.BenchTables(CK.REFERENCE_TABLES))
    x = M.gen_vector(coo.width, 7)
    gpu = CK.GpuCocktail(ctx, mat, x, np.zeros(coo.height, np.float32))
    out["parts"] = [p["kernel"] for p in gpu.part_infos()]
    n = 1000
    for flag, label in ((1, "do_spmv_graph_us"), (0, "do_spmv_direct_us")):
        L.check(L.lib().clspmv_set_option(b"cocktail_graph", flag), "set_option")
        gpu.do_spmv(64)
        ctx.sync()
        ctx.time_begin()
        gpu.do_spmv(n)
        out[label] = round(ctx.time_end() / n * 1e3, 3)
    k = 100

    def iterate(fn, label):
        for timed in (False, True):  # the first run captures any graphs
            gpu.result_from_cpu(np.zeros(coo.height, np.float32))
            gpu.vector_from_cpu(x)
            ctx.sync()
            t0 = time.perf_counter()
            fn(ctx, gpu, k)
            if timed:
                out[label] = round((time.perf_counter() - t0) / k * 1e6, 2)

    iterate(CK.power_iterate, "iterate_device_us_per_step")
    L.check(L.lib().clspmv_set_option(b"cocktail_graph", 1), "set_option")
    iterate(CK.power_iterate, "iterate_device_graph_us_per_step")
    L.check(L.lib().clspmv_set_option(b"cocktail_graph", 0), "set_option")
    iterate(CK.power_iterate_host, "iterate_host_loop_us_per_step")
    gpu.free()
    mat.free()
    coo.free()
    log(f"launch: {out}")
    return out


def load_traffic(kernel, layout_bytes):
    """HBM bytes per launch of `kernel` from the rocprofv3 counter runs
    summarised in profiles/ (FETCH_SIZE and WRITE_SIZE, each in a run of its
    own, scaled by what the counters read for a plain copy kernel).  The
    counters cannot be read inside this process, so the figure is the one
    recorded for the same kernel and layout; it is dropped (None) when the
    layout's bytes differ from those recorded with it (a changed layout)."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        t = json.load(f)
    v = t.get(kernel)
    if not isinstance(v, dict) or v.get("alg_bytes") != layout_bytes:
        return None, None
    src = (f"rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs of clspmv_amd.kprof on the same matrix and layout "
           f"({v.get('summary')}, measured {t.get('_calibration', {}).get('measured', 'with the summary')})")
    return v.get("hbm_bytes_per_launch"), src


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="c1", choices=["c1"])
    ap.add_argument("--scale", type=int, default=1, help="divide matrix sizes (testing only)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-c2", action="store_true", help="skip the Cocktail composite leg")
    ap.add_argument("--no-c3", action="store_true", help="skip the mixed-precision leg")
    ap.add_argument("--no-launch", action="store_true",
                    help="skip the launch-overhead leg (HIP graph capture), e.g. under a kernel tracer")
    ap.add_argument("--c4-matrices", type=int, default=1024,
                    help="matrices of the batched leg (default 1024, the c4 set; 0 skips the leg)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; --gpus > 1 is not supported "
                                   "(multi-GPU sharding is a design note only)"}))
        sys.exit(2)

    import numpy as np

    from clspmv_amd import matrix as M
    from clspmv_amd import workloads as W

    names = ["c1a", "c1b", "c1c", "c1d"]
    ctx = M.Context(0)
    work = []
    for n in names:
        t0 = time.perf_counter()
        wl = W.build(n, args.scale)
        part = M.Part(ctx, wl.fmt, wl.mat, wl.method)
        x = M.gen_vector(wl.coo.width, 7)
        dx = M.DeviceArray.from_numpy(ctx, x)
        dy = M.DeviceArray.from_numpy(ctx, np.zeros(wl.coo.height, np.float32))
        # copies of the part, x and y so that the per-kernel timing below can
        # rotate over a working set above 2 x the 256 MB Infinity Cache
        extra = []
        for _ in range(max(0, min(3, -(-600_000_000 // int(part.info["alg_bytes"])) - 1))):
            extra.append((M.Part(ctx, wl.fmt, wl.mat, wl.method), M.DeviceArray.from_numpy(ctx, x),
                          M.DeviceArray.from_numpy(ctx, np.zeros(wl.coo.height, np.float32))))
        # format arrays are on the device now; keep only the COO for the CPU leg
        wl.mat.free()
        # algorithmic bytes: the reference format's bytes for this matrix, or
        # the bytes this layout moves when that is less
        alg = min(W.reference_bytes(n, wl.coo), int(part.info["alg_bytes"]))
        work.append({"name": n, "coo": wl.coo, "x": x, "part": part, "dx": dx, "dy": dy, "desc": wl.describe,
                     "alg": alg, "ref_bytes": W.reference_bytes(n, wl.coo), "copies": extra})
        log(f"{n}: {wl.describe} rows={wl.coo.height} nnz={wl.coo.nnz} kernel={part.info['kernel']} "
            f"alg_bytes={part.info['alg_bytes']} ({time.perf_counter() - t0:.1f}s setup)")

    def step():
        for w in work:
            w["part"].spmv(w["dx"], w["dy"], 1)

    for _ in range(args.warmup):
        step()
    ctx.sync()
    ctx.time_begin()
    for _ in range(args.steps):
        step()
    ms_total = ctx.time_end()
    ctx.sync()
    ms_step = ms_total / args.steps
    flops = sum(2.0 * w["coo"].nnz for w in work)
    bytes_alg = sum(w["alg"] for w in work)
    gflops = flops / (ms_step * 1e-3) / 1e9
    gbps = bytes_alg / (ms_step * 1e-3) / 1e9
    log(f"step: {ms_step:.4f} ms  {gflops:.1f} GFLOP/s  {gbps:.1f} GB/s")

    # each kernel alone, HIP events on the context stream
    per = {}
    for w in work:
        k = max(10, args.steps)
        w["part"].spmv(w["dx"], w["dy"], 2)
        ctx.sync()
        ctx.time_begin()
        w["part"].spmv(w["dx"], w["dy"], k)
        ms_one = ctx.time_end() / k
        # the same launches rotating over the copies (working set > 512 MB),
        # so that no launch finds its matrix in the Infinity Cache
        rot = [(w["part"], w["dx"], w["dy"])] + w["copies"]
        ms = ms_one
        if len(rot) > 1:
            kr = k // len(rot) * len(rot)
            ctx.sync()
            ctx.time_begin()
            for i in range(kr):
                p_, x_, y_ = rot[i % len(rot)]
                p_.spmv(x_, y_, 1)
            ms = ctx.time_end() / kr
        info = w["part"].info
        alg = w["alg"]
        per[w["name"]] = {
            "format": w["desc"], "kernel": info["kernel"], "rows": info["rows"], "nnz": info["nnz"],
            "ms": round(ms, 5), "ms_single_copy": round(ms_one, 5), "rotated_copies": len(rot),
            "rotated_working_set_bytes": int(len(rot) * info["alg_bytes"]),
            "gflops": round(2.0 * info["nnz"] / (ms * 1e-3) / 1e9, 1),
            "alg_bytes": alg, "reference_format_bytes": w["ref_bytes"], "layout_bytes": info["alg_bytes"],
            "gbps": round(alg / (ms * 1e-3) / 1e9, 1),
            "pct_hbm_peak": round(100.0 * alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 1),
            # CSR-equivalent bytes (8 per nonzero + row pointers + x + y) and
            # stored bytes per nonzero of the format (its padding included)
            "csr_equiv_bytes": int(8 * info["nnz"] + 4 * (info["rows"] + 1) + 4 * info["cols"] + 8 * info["rows"]),
            "stored_bytes_per_nnz": round(info["stored_bytes"] / max(1, info["nnz"]), 3),
        }
        xs = w["part"].xtile_stats()
        if xs["tiles"]:
            per[w["name"]]["xtile_tiles"] = xs["tiles"]
            per[w["name"]]["xtile_staged_x_bytes_per_entry"] = round(xs["staged_x_bytes"] / max(1, info["nnz"]), 3)
        log(f"  {w['name']} {info['kernel']}: {ms * 1e3:.1f} us  {per[w['name']]['gbps']} GB/s "
            f"({per[w['name']]['pct_hbm_peak']}% of 8 TB/s)")
    # measured copy bandwidth (1 GiB device-to-device copy kernel): the
    # achievable streaming rate beside the published 8 TB/s peak
    import ctypes as C

    from clspmv_amd import _lib as L

    nb = 1 << 30
    src, dst = M.DeviceArray(ctx, nb // 4), M.DeviceArray(ctx, nb // 4)
    g = C.c_float()
    L.check(L.lib().clspmv_probe_copy(ctx.h, src.ptr, dst.ptr, nb, 10, C.byref(g)), "probe_copy")
    src.free()
    dst.free()
    copy_gbps = round(g.value, 1)
    log(f"copy kernel: {copy_gbps} GB/s (read + write of 1 GiB)")
    for n in per:
        per[n]["pct_of_copy"] = round(100.0 * per[n]["gbps"] / copy_gbps, 1)

    dom = max(per, key=lambda n: per[n]["ms"])
    traffic, traffic_src = load_traffic(per[dom]["kernel"], per[dom]["layout_bytes"])
    d = per[dom]
    achieved = d["alg_bytes"] / (d["ms"] * 1e-3) / 1e9
    roofline = {
        "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": traffic, "traffic_source": traffic_src,
        "kernel": d["kernel"], "matrix": dom, "alg_bytes_per_launch": d["alg_bytes"],
    }
    xs = next(w for w in work if w["name"] == dom)["part"].xtile_stats()
    if xs["tiles"]:
        # the x-tiled kernel's own bound: x staged through LDS per entry
        roofline["tiles_per_spmv"] = xs["tiles"]
        roofline["staged_x_bytes_per_spmv"] = xs["staged_x_bytes"]
        roofline["staged_x_bytes_per_entry"] = round(xs["staged_x_bytes"] / max(1, d["nnz"]), 3)

    c3 = None if args.no_c3 else mixed_c3(ctx, args, work, per)
    c2 = None if args.no_c2 else cocktail_c2(ctx, args)
    c4 = batched_c4(ctx, args) if args.c4_matrices > 0 else None
    launch = None if args.no_launch else launch_c0(ctx)

    cpu = None
    if not args.no_cpu:
        log("timing the reference CPU SpMV (1 core)...")
        cpu = cpu_baseline(work)
        log(f"cpu: {cpu['value']} GFLOP/s ({cpu['sample']}); project OpenMP CSR "
            f"{cpu['project_omp']['value']} GFLOP/s on {cpu['project_omp']['cores']} threads")

    for w in work:
        w["part"].free()
        w["dx"].free()
        w["dy"].free()
        for p_, x_, y_ in w["copies"]:
            p_.free()
            x_.free()
            y_.free()
    ctx.close()

    out = {
        "metric": METRIC,
        "value": round(gflops, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generators, resident in HBM)",
        "config": {
            "workload": "c1: one FP32 y+=Ax on each of c1a 27-pt stencil 128^3, c1b 4M x 32 uniform, "
                        "c1c 2M Zipf(1.8) mean 16 max 50k, c1d 1M block rows 3x3 x 9; formats: "
                        + ", ".join(f"{w['name']} {w['desc']}" for w in work)
                        + (f" [sizes / {args.scale}]" if args.scale > 1 else ""),
            "matrices": names,
            "nnz_total": int(sum(w["coo"].nnz for w in work)),
            "bytes_alg_per_step": int(bytes_alg),
        },
        "hbm_gbps": round(gbps, 1),
        "copy_gbps": copy_gbps,
        "pct_hbm_peak": round(100.0 * gbps / HBM_PEAK_GBPS, 2),
        "per_matrix": per,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "mixed_c3": c3,
        "cocktail_c2": c2,
        "batched_c4": c4,
        "launch_c0": launch,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
