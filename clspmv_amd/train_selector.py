"""Trains the learned format selector on autotuner-timed matrices and
reports its regret against exhaustive autotuning on held-out matrices.

    python -m clspmv_amd.train_selector --n 2400 --log2-min 10 --log2-max 20 --out DIR

DIR receives dataset.npz (features, candidate times, kinds), model.npz (the
trained network and its feature normalisation) and report.json: per split the
mean / median / p90 regret of the network's choice, of the fixed default
(CSR adaptive row blocks) and -- on a subset of the held-out matrices -- of
the Cocktail partition made with the MI355X tables, plus the share of
matrices within 10 % of the best candidate.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import _lib as L
from . import cocktail as CK
from . import matrix as M
from . import selector as S


def log(s):
    print(s, file=sys.stderr, flush=True)


def summary(r):
    r = np.asarray(r, np.float64)
    return {"mean": float(np.mean(r)), "median": float(np.median(r)), "p90": float(np.percentile(r, 90)),
            "within_10pct": float(np.mean(r <= 0.10)), "n": int(len(r))}


def cocktail_times(ctx, idx, seed, log2_min, log2_max, detail=None, target_ms=1.0):
    """Cocktail (MI355X tables, the analyzer's plan) time of matrices `idx`
    of the dataset family, timed like the autotuner's candidates.  With
    `detail` (a dict), also records per matrix its kind, size and kernels."""
    tables = CK.BenchTables(CK.MI355X_TABLES)
    rng = np.random.default_rng(seed)
    out = {}
    want = set(int(i) for i in idx)
    for i in range(max(want) + 1):
        kind, coo = S.random_matrix(i, rng, log2_min, log2_max)
        if i in want:
            mat = CK.analyze_matrix(coo.copy(), tables)
            gpu = CK.GpuCocktail(ctx, mat, M.gen_vector(coo.width, 1), np.zeros(coo.height, np.float32))
            gpu.do_spmv(1)
            ctx.sync()
            ctx.time_begin()
            gpu.do_spmv(2)
            ms = ctx.time_end() / 2
            n = int(min(50, max(3, target_ms / max(ms, 1e-4))))
            gpu.do_spmv(n)  # untimed, as the autotuner does for each candidate
            ctx.time_begin()
            gpu.do_spmv(n)
            out[i] = ctx.time_end() / n
            if detail is not None:
                detail[i] = {"kind": kind, "rows": int(coo.height), "nnz": int(coo.nnz),
                             "kernels": [p["kernel"] for p in gpu.part_infos()], "ms": out[i]}
            gpu.free()
            mat.free()
        coo.free()
    return out


def eval_cocktail(ctx, data_dir, n, subset, seed, log2_min, log2_max, test_frac=0.2, detail=None):
    """Regret of the Cocktail plan (MI355X tables) on the first `subset`
    held-out matrices of a saved dataset, against its autotuned times.
    With `detail` (a list), appends one record per matrix: its plan's
    kernels and time next to the best candidate's."""
    d = np.load(os.path.join(data_dir, "dataset.npz"))
    times = d["times"]
    names = [str(c) for c in d["candidates"]]
    rng = np.random.default_rng(seed + 1)
    perm = rng.permutation(n)
    test = np.sort(perm[:int(n * test_frac)])[:subset]
    rec = {}
    ck = cocktail_times(ctx, test, seed, log2_min, log2_max, rec)
    r = np.array([ck[int(i)] / times[i].min() - 1.0 for i in test])
    if detail is not None:
        for i in test:
            b = int(np.argmin(times[i]))
            detail.append(dict(rec[int(i)], index=int(i), best=names[b], best_ms=float(times[i][b]),
                               regret=float(ck[int(i)] / times[i][b] - 1.0)))
    kinds = d["kinds"][test]
    out = {"all": summary(r)}
    for k in sorted(set(kinds.tolist())):
        out[k] = summary(r[kinds == k])
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--eval-cocktail", metavar="DIR", default=None,
                    help="only re-time the Cocktail plan on the held-out subset of the dataset in DIR")
    ap.add_argument("--n", type=int, default=2400)
    ap.add_argument("--test-frac", type=float, default=0.2)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--log2-min", type=float, default=10)
    ap.add_argument("--log2-max", type=float, default=20)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--cocktail-subset", type=int, default=200)
    ap.add_argument("--out", default=None)
    ap.add_argument("--detail", default=None, help="write the Cocktail plan's per-matrix records (kernels, time, best candidate) to this JSON file")
    ap.add_argument("--option", action="append", default=[], metavar="NAME=VALUE",
                    help="clspmv_set_option before timing (e.g. cocktail_streams=0); repeatable")
    args = ap.parse_args(argv)
    for opt in args.option:
        name, value = opt.split("=", 1)
        L.check(L.lib().clspmv_set_option(name.encode(), int(value)), "set_option")
    ctx = M.Context(0)
    if args.eval_cocktail:
        detail = [] if args.detail else None
        res = eval_cocktail(ctx, args.eval_cocktail, args.n, args.cocktail_subset, args.seed, args.log2_min,
                            args.log2_max, args.test_frac, detail)
        if args.detail:
            with open(args.detail, "w") as f:
                json.dump(detail, f, indent=0)
        print(json.dumps({"cocktail_mi355x_tables": res}, indent=1), flush=True)
        ctx.close()
        return
    if not args.out:
        ap.error("--out is required")
    os.makedirs(args.out, exist_ok=True)
    t0 = time.perf_counter()
    feats, times, kinds, nnz = S.build_dataset(ctx, args.n, args.seed, log, args.log2_min, args.log2_max)
    np.savez(os.path.join(args.out, "dataset.npz"), features=feats, times=times, kinds=kinds, nnz=nnz,
             candidates=np.array(S.CANDIDATE_NAMES))
    log(f"[selector] dataset: {args.n} matrices in {time.perf_counter() - t0:.0f} s")
    rng = np.random.default_rng(args.seed + 1)
    perm = rng.permutation(args.n)
    n_test = int(args.n * args.test_frac)
    test, train = np.sort(perm[:n_test]), np.sort(perm[n_test:])
    t1 = time.perf_counter()
    sel = S.Selector.fit(ctx, feats[train], times[train], steps=args.steps)
    log(f"[selector] trained in {time.perf_counter() - t1:.1f} s")
    sel.save(os.path.join(args.out, "model.npz"))
    report = {"n_train": int(len(train)), "n_test": int(len(test)), "candidates": S.CANDIDATE_NAMES,
              "features": S.FEATURE_NAMES}
    for name, idx in (("train", train), ("test", test)):
        t = times[idx]
        pick = sel.predict(feats[idx], t)
        default = np.full(len(idx), S.DEFAULT)
        report[name] = {
            "network": summary(S.regret(t, pick)),
            "default_csr_adaptive": summary(S.regret(t, default)),
            "accuracy": float(np.mean(pick == np.argmin(t, axis=1))),
            "best_counts": {S.CANDIDATE_NAMES[k]: int(v) for k, v in
                            zip(*np.unique(np.argmin(t, axis=1), return_counts=True))},
        }
    sub = test[:args.cocktail_subset]
    if len(sub):
        rec = {}
        ck = cocktail_times(ctx, sub, args.seed, args.log2_min, args.log2_max, rec)
        r = np.array([ck[int(i)] / times[i].min() - 1.0 for i in sub])
        report["test"]["cocktail_mi355x_tables"] = summary(r)
        report["test"]["cocktail_mi355x_tables_by_kind"] = {
            k: summary(r[kinds[sub] == k]) for k in sorted(set(kinds[sub].tolist()))}
        if args.detail:
            with open(args.detail, "w") as f:
                json.dump([dict(rec[int(i)], index=int(i), best=S.CANDIDATE_NAMES[int(np.argmin(times[i]))],
                                best_ms=float(times[i].min()), regret=float(ck[int(i)] / times[i].min() - 1.0))
                           for i in sub], f, indent=0)
        pick = sel.predict(feats[sub], times[sub])
        report["test"]["network_on_cocktail_subset"] = summary(S.regret(times[sub], pick))
    report["seconds"] = round(time.perf_counter() - t0, 1)
    with open(os.path.join(args.out, "report.json"), "w") as f:
        json.dump(report, f, indent=1)
    print(json.dumps({k: report[k] for k in ("train", "test")}, indent=1), flush=True)
    sel.net.free()
    ctx.close()


if __name__ == "__main__":
    main()
