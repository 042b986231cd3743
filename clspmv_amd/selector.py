"""Learned format selector and the exhaustive autotuner it learns from.

The autotuner times every applicable (format, kernel) candidate of this
library on a matrix and returns the times; the selector is a small network
(include/clspmv_mlp.h: 24 structural features -> 2 x 128 ReLU -> one class
per candidate), trained on the GPU by the library's own kernels on
autotuner-timed matrices, that predicts the fastest candidate from the
features alone.  Exhaustive autotuning stays available as its fallback and
as the measure of its regret (time of the predicted candidate / time of the
best one - 1).

    times = autotune(ctx, coo)                     # one row of timings
    sel = Selector.fit(ctx, feats, times)          # GPU training
    fmt, method, name = sel.choose(coo)
"""
from __future__ import annotations

import ctypes as C
import math
import os
import time

import numpy as np

from . import _lib as L
from . import matrix as M
from ._lib import check, lib

MODEL_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tables", "selector")

# ------------------------------------------------------------------ network


class Mlp:
    """ctypes wrapper of clspmv_mlp_*."""

    def __init__(self, ctx, d, h, c, seed=0):
        self.ctx, self.d, self.h, self.c = ctx, d, h, c
        self.hdl = C.c_void_p()
        check(lib().clspmv_mlp_create(ctx.h, d, h, c, seed, C.byref(self.hdl)), "mlp_create")
        self.np = lib().clspmv_mlp_num_params(self.hdl)

    def get_params(self):
        p = np.empty(self.np, np.float32)
        check(lib().clspmv_mlp_get_params(self.hdl, p.ctypes.data_as(C.POINTER(C.c_float))), "mlp_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        check(lib().clspmv_mlp_set_params(self.hdl, p.ctypes.data_as(C.POINTER(C.c_float))), "mlp_set_params")

    def loss_grad(self, x, y):
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.int32)
        loss = C.c_float()
        g = np.empty(self.np, np.float32)
        check(lib().clspmv_mlp_loss_grad(self.hdl, x.ctypes.data_as(C.POINTER(C.c_float)),
                                         y.ctypes.data_as(C.POINTER(C.c_int)), len(y), C.byref(loss),
                                         g.ctypes.data_as(C.POINTER(C.c_float))), "mlp_loss_grad")
        return loss.value, g

    def train(self, x, y, steps=2000, lr=2e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.int32)
        loss = C.c_float()
        check(lib().clspmv_mlp_train(self.hdl, x.ctypes.data_as(C.POINTER(C.c_float)),
                                     y.ctypes.data_as(C.POINTER(C.c_int)), len(y), steps, lr, beta1, beta2, eps,
                                     C.byref(loss)), "mlp_train")
        return loss.value

    def predict(self, x):
        x = np.ascontiguousarray(x, np.float32)
        n = x.shape[0]
        z = np.empty((n, self.c), np.float32)
        cls = np.empty(n, np.int32)
        check(lib().clspmv_mlp_predict(self.hdl, x.ctypes.data_as(C.POINTER(C.c_float)), n,
                                       z.ctypes.data_as(C.POINTER(C.c_float)),
                                       cls.ctypes.data_as(C.POINTER(C.c_int))), "mlp_predict")
        return z, cls

    def free(self):
        if getattr(self, "hdl", None) and self.ctx.h:
            lib().clspmv_mlp_free(self.hdl)
        self.hdl = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ------------------------------------------------------------------ features

FEATURE_NAMES = [
    "log2_rows", "log2_cols", "log2_nnz", "mean_row", "row_cv", "log2_max_over_mean", "empty_rows",
    "rows_over_64", "nnz_in_rows_over_64", "log2_ell_padding", "log2_diagonals", "nnz_top7_diagonals",
    "nnz_top27_diagonals", "log2_dia_padding", "mean_abs_offset_over_cols", "nnz_within_1024_of_diagonal",
    "block_fill_2", "block_fill_3", "block_fill_4", "block_fill_8", "log2_x_over_l2", "segments16_per_nnz",
    "adjacent_column_pairs", "p90_row_over_mean",
]


SAMPLE_NNZ = 1 << 20  # entries above which the locality features use a row sample


def features(coo) -> np.ndarray:
    """24 structural features of a COO matrix (rows sorted).  Row-length
    statistics use every row; the diagonal, block, segment and adjacency
    features of a matrix with more than SAMPLE_NNZ entries are computed on
    every s-th window of 48 rows (48 is a multiple of every block height),
    s chosen so that about SAMPLE_NNZ entries remain."""
    r = coo.row_id.astype(np.int64)
    c = coo.col_id.astype(np.int64)
    R, W, n_all = coo.height, coo.width, coo.nnz
    lens = np.bincount(r, minlength=R).astype(np.float64)
    mean = n_all / max(R, 1)
    n = n_all
    f = [math.log2(max(R, 1)), math.log2(max(W, 1)), math.log2(max(n_all, 1)), mean,
         float(lens.std() / max(mean, 1e-9)), math.log2(max(lens.max(initial=0) / max(mean, 1e-9), 1e-9) + 1e-9)
         if n_all else 0.0, float(np.mean(lens == 0)), float(np.mean(lens > 64)),
         float(lens[lens > 64].sum() / max(n_all, 1)), math.log2(max(lens.max(initial=0) * R / max(n_all, 1), 1.0))]
    if n_all > SAMPLE_NNZ:
        stride = -(-n_all // SAMPLE_NNZ)
        keep = (r // 48) % stride == 0
        r, c = r[keep], c[keep]
        n = len(r)
    if n:
        d = c - r
        _, cnt = np.unique(d, return_counts=True)
        cnt = np.sort(cnt)[::-1]
        f += [math.log2(len(cnt) + 1), float(cnt[:7].sum() / n), float(cnt[:27].sum() / n),
              math.log2(max(len(cnt) * R / n_all, 1e-9)), float(np.mean(np.abs(d)) / max(W, 1)),
              float(np.mean(np.abs(d) <= 1024))]
        for b in (2, 3, 4, 8):
            nb = len(np.unique((r // b) * (W // b + 1) + c // b))
            f.append(float(n / (nb * b * b)))
        f.append(math.log2(4.0 * W / (4 << 20)))
        f.append(float(len(np.unique(r * (W // 16 + 1) + c // 16)) / n))
        same = (r[1:] == r[:-1])
        f.append(float(np.sum(same & (c[1:] == c[:-1] + 1)) / n))
        f.append(float(np.percentile(lens, 90) / max(mean, 1e-9)))
    else:
        f += [0.0] * 6 + [0.0] * 4 + [math.log2(4.0 * max(W, 1) / (4 << 20)), 0.0, 0.0, 0.0]
    assert len(f) == len(FEATURE_NAMES)
    return np.array(f, np.float64)


# ------------------------------------------------------------------ candidates

# (name, format, method, conversion key)
CANDIDATES = [
    ("dia_row", L.FMT_DIA, 0, "dia"),
    ("dia_row4", L.FMT_DIA, 1, "dia"),
    ("bdia_lds", L.FMT_BDIA, 1, "bdia"),
    ("ell_row4", L.FMT_ELL, 1, "ell"),
    ("ell_xtile", L.FMT_ELL, 4, "ell"),
    ("csr_adaptive", L.FMT_CSR_VECTOR, 1, "csr"),
    ("csr_sliced", L.FMT_CSR_VECTOR, 2, "csr"),
    ("csr_xtile", L.FMT_CSR_VECTOR, 4, "csr"),
    ("coo_segment", L.FMT_COO, 0, "coo"),
    ("bsr2_sliced", L.FMT_BSR, 0, "bsr2"),
    ("bsr3_sliced", L.FMT_BSR, 0, "bsr3"),
    ("bsr4_sliced", L.FMT_BSR, 0, "bsr4"),
    # the reference's own tuning sweeps (A/src/bench_sell.cpp: slice heights
    # 32 and 128; A/src/bench_sbell.cpp, bench_bcsr.cpp: block shapes)
    ("sell32", L.FMT_SELL, 0, "sell32"),
    ("sell128", L.FMT_SELL, 0, "sell128"),
    ("sbell_1x4", L.FMT_SBELL, 0, "sbell14"),
    ("sbell_2x4", L.FMT_SBELL, 0, "sbell24"),
    ("sbell_4x4", L.FMT_SBELL, 0, "sbell44"),
    ("bcsr_4x4", L.FMT_BCSR, 0, "bcsr44"),
    # block-CSR with 2-16 lanes per block row (by the mean blocks per row):
    # more waves than the sliced kernel's one lane per block row
    ("bsr2_lanes", L.FMT_BSR, 1, "bsr2"),
    ("bsr3_lanes", L.FMT_BSR, 1, "bsr3"),
    ("bsr4_lanes", L.FMT_BSR, 1, "bsr4"),
    # DIA with 2-8 lanes per row splitting the diagonals (small matrices)
    ("dia_lanes", L.FMT_DIA, 4, "dia"),
]
CANDIDATE_NAMES = [c[0] for c in CANDIDATES]
DEFAULT = CANDIDATE_NAMES.index("csr_adaptive")


def applicable(feats):
    """Candidates worth converting to, from the features alone: DIA/BDIA when
    the diagonal padding stays below 2x the nonzeros (coo2dia refuses at 2x),
    ELL when its padding stays below 4x, BSR b x b when the blocks are at
    least half full, SELL when ELL would pad less than 64x, the reference's
    SBELL / BCSR block shapes when 4 x 4 blocks are at least half full.
    Row-list kernels always apply (x-tiled needs >= 4 columns)."""
    f = np.atleast_2d(feats)
    ok = np.ones((f.shape[0], len(CANDIDATES)), bool)
    for i, (_, _, _, key) in enumerate(CANDIDATES):
        if key in ("dia", "bdia"):
            ok[:, i] = f[:, 13] < 1.0
        elif key == "ell":
            ok[:, i] = f[:, 9] <= 2.0
        elif key.startswith("bsr"):
            ok[:, i] = f[:, 16 + (2, 3, 4).index(int(key[3]))] >= 0.5
        elif key.startswith("sell"):
            ok[:, i] = f[:, 9] <= 6.0
        elif key.startswith(("sbell", "bcsr")):
            ok[:, i] = f[:, 18] >= 0.5
    xt = [i for i, c in enumerate(CANDIDATES) if c[2] == 4]
    ok[:, xt] &= (f[:, 1:2] >= 2.0)  # log2(cols) >= 2
    return ok


def _convert(key, coo):
    """The candidate's input format, or None where the conversion refuses."""
    try:
        if key == "coo":
            return coo
        if key == "csr":
            return M.coo2csr(coo)
        if key == "dia":
            return M.coo2dia(coo)
        if key == "bdia":
            return M.coo2bdia(coo)
        if key == "ell":
            return M.coo2ell(coo)
        if key.startswith("bsr"):
            b = int(key[3])
            return M.coo2bsr(coo, b, b)
        if key.startswith("sell"):
            return M.coo2sell(coo, int(key[4:]))
        if key.startswith("sbell"):
            return M.coo2sbell(coo, int(key[6]), int(key[5]), 32)
        if key.startswith("bcsr"):
            return M.coo2b4csr(coo, int(key[5]), int(key[4]))
    except L.ClspmvError:
        return None
    raise KeyError(key)


def autotune(ctx, coo, feats=None, target_ms=1.0):
    """Time every applicable candidate (ms per y += A x); inf where a
    candidate does not apply."""
    feats = features(coo) if feats is None else feats
    ok = applicable(feats)[0]
    x = M.DeviceArray.from_numpy(ctx, M.gen_vector(coo.width, 1))
    y = M.DeviceArray.from_numpy(ctx, np.zeros(coo.height, np.float32))
    times = np.full(len(CANDIDATES), np.inf)
    cache = {}
    try:
        for i, (name, fmt, meth, key) in enumerate(CANDIDATES):
            if not ok[i]:
                continue
            if key not in cache:
                cache[key] = _convert(key, coo)
            mat = cache[key]
            if mat is None:
                continue
            try:
                part = M.Part(ctx, fmt, mat, meth)
            except L.ClspmvError:
                continue
            part.spmv(x, y, 1)
            ctx.sync()
            ctx.time_begin()
            part.spmv(x, y, 2)
            ms = ctx.time_end() / 2
            n = int(min(50, max(3, target_ms / max(ms, 1e-4))))
            part.spmv(x, y, n)  # untimed: the GPU at its running clock, as in the timed loop
            ctx.time_begin()
            part.spmv(x, y, n)
            times[i] = ctx.time_end() / n
            part.free()
    finally:
        for k, mat in cache.items():
            if mat is not None and mat is not coo:
                mat.free()
        x.free()
        y.free()
    return times


# ------------------------------------------------------------------ matrices

KINDS = ["stencil27", "lap2d", "uniform", "zipf", "blocks", "cocktail"]


def random_matrix(i, rng, log2_min=12, log2_max=18):
    """Matrix i of the training / evaluation family: one of the benchmark
    generators (27-point stencil, 5-point Laplacian, uniform rows, Zipf rows,
    dense blocks, Cocktail composite) at a log-uniform row count."""
    kind = KINDS[i % len(KINDS)]
    rows = int(2 ** rng.uniform(log2_min, log2_max))
    seed = 1000 + i
    if kind == "stencil27":
        return kind, M.gen_stencil27(max(4, round(rows ** (1 / 3))), seed)
    if kind == "lap2d":
        s = max(8, round(math.sqrt(rows)))
        return kind, M.gen_laplacian2d(s, s)
    if kind == "uniform":
        return kind, M.gen_uniform_rows(rows, int(rng.choice([2, 4, 8, 16, 32, 48])), seed)
    if kind == "zipf":
        return kind, M.gen_zipf_rows(rows, float(rng.uniform(1.5, 2.3)), float(rng.choice([4.0, 8.0, 16.0, 32.0])),
                                     min(50000, rows // 2), seed)
    if kind == "blocks":
        b = int(rng.choice([2, 3, 4]))
        return kind, M.gen_block_rows(max(16, rows // b), b, int(rng.integers(2, 13)), int(rng.choice([8, 64, 512, 5000])),
                                      seed)
    return kind, M.gen_cocktail_composite(max(1024, rows // 16 * 16), seed)  # rows a multiple of 16


def build_dataset(ctx, n, seed=0, log=None, log2_min=12, log2_max=18):
    """Features and autotuner times of n random matrices."""
    rng = np.random.default_rng(seed)
    feats, times, kinds, nnz = [], [], [], []
    t0 = time.perf_counter()
    for i in range(n):
        kind, coo = random_matrix(i, rng, log2_min, log2_max)
        f = features(coo)
        feats.append(f)
        times.append(autotune(ctx, coo, f))
        kinds.append(kind)
        nnz.append(coo.nnz)
        coo.free()
        if log and (i + 1) % 100 == 0:
            log(f"[selector] {i + 1}/{n} matrices timed ({time.perf_counter() - t0:.0f} s)")
    return np.array(feats), np.array(times), np.array(kinds), np.array(nnz)


# ------------------------------------------------------------------ selector


def regret(times, choice):
    """Per matrix: time of the chosen candidate / best time - 1."""
    best = times.min(axis=1)
    t = times[np.arange(len(choice)), choice]
    return t / best - 1.0


class Selector:
    """Feature normalisation + the trained network."""

    def __init__(self, net, mu, sd):
        self.net, self.mu, self.sd = net, mu, sd

    @classmethod
    def fit(cls, ctx, feats, times, hidden=128, steps=3000, lr=2e-3, seed=7):
        ok = np.isfinite(times).any(axis=1)
        feats, times = feats[ok], times[ok]
        labels = np.argmin(times, axis=1).astype(np.int32)
        mu = feats.mean(axis=0)
        sd = feats.std(axis=0) + 1e-6
        net = Mlp(ctx, feats.shape[1], hidden, len(CANDIDATES), seed)
        net.train(((feats - mu) / sd).astype(np.float32), labels, steps=steps, lr=lr)
        return cls(net, mu, sd)

    def predict(self, feats, times=None):
        """Candidate index per matrix: the largest logit among the candidates
        that apply (from the features; with `times`, also excluding those the
        autotuner could not build)."""
        f = np.atleast_2d(feats)
        z, _ = self.net.predict(((f - self.mu) / self.sd).astype(np.float32))
        mask = applicable(f)
        if times is not None:
            mask &= np.isfinite(np.atleast_2d(times))
        return np.argmax(np.where(mask, z, -np.inf), axis=1)

    def choose(self, coo):
        f = features(coo)
        i = int(self.predict(f)[0])
        name, fmt, meth, _ = CANDIDATES[i]
        return fmt, meth, name

    def save(self, path):
        np.savez(path, params=self.net.get_params(), mu=self.mu, sd=self.sd, hidden=self.net.h)

    @classmethod
    def load_default(cls, ctx):
        """The network trained on MI355X timings (tables/selector/model.npz)."""
        return cls.load(ctx, os.path.join(MODEL_DIR, "model.npz"))

    @classmethod
    def load(cls, ctx, path):
        z = np.load(path, allow_pickle=False)
        net = Mlp(ctx, len(z["mu"]), int(z["hidden"]), len(CANDIDATES))
        net.set_params(z["params"])
        return cls(net, z["mu"], z["sd"])
