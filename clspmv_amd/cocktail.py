"""Host mirror of the Cocktail API (A/include/analyze.h, A/include/spmv_cocktail.h).

    tables = BenchTables(REFERENCE_TABLES)          # ReadBench
    mat = analyze_matrix(coo, tables)               # analyze_matrix
    gpu = GpuCocktail(ctx, mat, vec, res)           # init_mat_kernels
    gpu.do_spmv(1)                                  # do_spmv
    y = gpu.result_to_cpu()                         # cpy_result_to_cpu

All work runs in libclspmv (C++ analyzer, HIP kernels).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib as L
from ._lib import check, lib
from .matrix import Context

HERE = os.path.dirname(os.path.abspath(__file__))
REFERENCE_TABLES = os.path.join(HERE, "tables", "reference_amd")
MI355X_TABLES = os.path.join(HERE, "tables", "mi355x")
PART_NAMES = ("bdia", "dia", "sbell", "bell", "bcsr", "sell", "ell", "csr", "coo")


class BenchTables:
    """Per-format performance tables (benchmark_all) read from a directory of
    .ben files in the reference's schema."""

    def __init__(self, path: str = REFERENCE_TABLES):
        self.path = path
        self.h = C.c_void_p()
        check(lib().clspmv_bench_load(path.encode(), C.byref(self.h)), f"bench_load({path})")

    def maxflop(self, rows: int, nnz: int) -> L.MaxflopInfo:
        m = L.MaxflopInfo()
        check(lib().clspmv_maxflop(self.h, rows, nnz, C.byref(m)), "maxflop")
        return m

    def overhead_us(self, groups: int, threads: int) -> float:
        return lib().clspmv_find_overhead(self.h, groups, threads)

    def __del__(self):
        try:
            if self.h:
                lib().clspmv_bench_free(self.h)
        except Exception:
            pass


class CocktailMatrix:
    """Owns a cocktail<int,int,float> partition."""

    def __init__(self):
        self.s = L.Cocktail()
        lib().clspmv_init_cocktail(C.byref(self.s))

    def parts(self):
        return [n for n in PART_NAMES if getattr(self.s, f"ifuse{n}")]

    def part_nnz(self):
        return {n: getattr(self.s, n).matinfo.nnz for n in self.parts()}

    def describe(self):
        d = {"width": self.s.mat_width, "height": self.s.mat_height, "nnz": self.s.mat_nnz, "parts": {}}
        for n in self.parts():
            d["parts"][n] = {"nnz": getattr(self.s, n).matinfo.nnz, "method": getattr(self.s, f"{n}_meth_num")}
        if self.s.ifusecsr:
            d["parts"]["csr"]["groups"] = self.s.csr_group_num
        if self.s.ifusecoo:
            d["parts"]["coo"]["groups"] = self.s.coo_group_num
        return d

    def free(self):
        if getattr(self, "s", None) is not None:
            lib().clspmv_free_cocktail(C.byref(self.s))
            self.s = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def analyze_matrix(coo, tables: BenchTables, dia_priority_multiplier: float = 1.0,
                   estimate_flat_partial: bool = True, estimate_flat_full: bool = False,
                   verbose: bool = False) -> CocktailMatrix:
    """analyze_matrix (A/include/analyze.h:257); sorts an unsorted `coo`."""
    mat = CocktailMatrix()
    check(lib().clspmv_analyze_matrix(tables.h, C.byref(coo.s), C.byref(mat.s), dia_priority_multiplier,
                                      int(estimate_flat_partial), int(estimate_flat_full), int(verbose)),
          "analyze_matrix")
    return mat


FORMAT_NAMES = {L.FMT_CSR_SCALAR: "csr_scalar", L.FMT_CSR_VECTOR: "csr", L.FMT_BDIA: "bdia", L.FMT_DIA: "dia",
                L.FMT_ELL: "ell", L.FMT_COO: "coo", L.FMT_BELL: "bell", L.FMT_BCSR: "bcsr", L.FMT_SELL: "sell",
                L.FMT_SBELL: "sbell", L.FMT_BSR: "bsr"}


class Plan:
    """The MI355X plan of a partition (clspmv_analyze_plan /
    clspmv_plan_cocktail): the parts that run, each with its format, kernel
    method and row range."""

    def __init__(self, h, mat=None):
        self.h = h
        self._mat = mat  # keeps a caller's partition alive

    @classmethod
    def analyze(cls, coo, tables: BenchTables, dia_priority_multiplier: float = 1.0,
                estimate_flat_partial: bool = True, estimate_flat_full: bool = False):
        h = C.c_void_p()
        check(lib().clspmv_analyze_plan(tables.h, C.byref(coo.s), dia_priority_multiplier,
                                        int(estimate_flat_partial), int(estimate_flat_full), C.byref(h)),
              "analyze_plan")
        return cls(h)

    @classmethod
    def of(cls, mat: CocktailMatrix):
        h = C.c_void_p()
        check(lib().clspmv_plan_cocktail(C.byref(mat.s), C.byref(h)), "plan_cocktail")
        return cls(h, mat)

    def parts(self):
        out = []
        for i in range(lib().clspmv_plan_num_parts(self.h)):
            q = L.PlanPart()
            check(lib().clspmv_plan_get_part(self.h, i, C.byref(q)), "plan_get_part")
            d = {k: getattr(q, k) for k, _ in L.PlanPart._fields_}
            d["source_name"] = PART_NAMES[q.source] if q.source < 9 else "run"
            d["format_name"] = FORMAT_NAMES.get(q.format, str(q.format))
            out.append(d)
        return out

    def partition(self):
        """The partition struct the plan runs (a read-only view)."""
        p = lib().clspmv_plan_partition(self.h)
        return C.cast(p, C.POINTER(L.Cocktail)).contents

    def free(self):
        if getattr(self, "h", None):
            lib().clspmv_plan_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class GpuCocktail:
    """cocktail_gpu + cocktail_kernels: the partition on the device."""

    def __init__(self, ctx: Context, mat, vec=None, res=None, use_tex: bool = False):
        """`mat`: a CocktailMatrix (init_mat_kernels) or a Plan
        (init_plan_kernels)."""
        self.ctx, self.mat = ctx, mat
        fp = C.POINTER(C.c_float)
        v = None if vec is None else np.ascontiguousarray(vec, np.float32)
        r = None if res is None else np.ascontiguousarray(res, np.float32)
        self.h = C.c_void_p()
        vp = None if v is None else v.ctypes.data_as(fp)
        rp = None if r is None else r.ctypes.data_as(fp)
        if isinstance(mat, Plan):
            s = mat.partition()
            self._shape = (s.mat_width, s.mat_height)
            check(lib().clspmv_init_plan_kernels(ctx.h, mat.h, vp, rp, C.byref(self.h)), "init_plan_kernels")
        else:
            self._shape = (mat.s.mat_width, mat.s.mat_height)
            check(lib().clspmv_init_mat_kernels(ctx.h, C.byref(mat.s), vp, rp, int(use_tex), C.byref(self.h)),
                  "init_mat_kernels")

    @property
    def width(self):
        return self._shape[0]

    @property
    def height(self):
        return self._shape[1]

    def do_spmv(self, ntimes: int = 1):
        check(lib().clspmv_do_spmv(self.h, ntimes), "do_spmv")

    def vector_from_cpu(self, vec):
        v = np.ascontiguousarray(vec, np.float32)
        check(lib().clspmv_cpy_vector_from_cpu(self.h, v.ctypes.data_as(C.POINTER(C.c_float)), v.size), "cpy_vec")

    def vector_from_gpu(self, dptr, n):
        check(lib().clspmv_cpy_vector_from_gpu(self.h, dptr, n), "cpy_vec_gpu")

    def result_from_cpu(self, res):
        r = np.ascontiguousarray(res, np.float32)
        check(lib().clspmv_cpy_result_from_cpu(self.h, r.ctypes.data_as(C.POINTER(C.c_float)), r.size), "cpy_res")

    def result_to_cpu(self):
        out = np.empty(self.height, np.float32)
        check(lib().clspmv_cpy_result_to_cpu(self.h, out.ctypes.data_as(C.POINTER(C.c_float)), out.size),
              "cpy_res_to_cpu")
        return out

    def buffers(self):
        x, y = C.c_void_p(), C.c_void_p()
        check(lib().clspmv_gpumat_buffers(self.h, C.byref(x), C.byref(y)), "buffers")
        return x, y

    def part_infos(self):
        out = []
        for i in range(lib().clspmv_gpumat_num_parts(self.h)):
            inf = L.PartInfo()
            check(lib().clspmv_gpumat_part_info(self.h, i, C.byref(inf)), "part_info")
            d = {k: getattr(inf, k) for k, _ in L.PartInfo._fields_}
            d["kernel"] = inf.kernel.decode()
            out.append(d)
        return out

    def free(self):
        if getattr(self, "h", None) and self.ctx.h:
            lib().clspmv_free_mat_kernels(self.h)
        self.h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def power_iterate(ctx: Context, gpu: GpuCocktail, k: int):
    """The iteration of A/src/eval.cpp:19-72 (`example`): k times
    {y += A x; x := y}, device-resident (clspmv_power_iterate: captured
    graphs, no host round trip); returns y."""
    check(lib().clspmv_power_iterate(gpu.h, k), "power_iterate")
    return gpu.result_to_cpu()


def power_iterate_host(ctx: Context, gpu: GpuCocktail, k: int):
    """The same iteration driven from the host as the reference's example
    does it: one do_spmv, then cpy_vector_from_gpu(y), k times."""
    _, dy = gpu.buffers()
    for _ in range(k):
        gpu.do_spmv(1)
        gpu.vector_from_gpu(dy, gpu.width)
    return gpu.result_to_cpu()


class EvalReport(C.Structure):
    _fields_ = [("distance", C.c_double), ("max_diff", C.c_double), ("max_diff_ratio", C.c_double),
                ("max_rel_err", C.c_double), ("max_diff_id", C.c_int), ("ntimes", C.c_int), ("nparts", C.c_int),
                ("ms_per_spmv", C.c_double), ("gflops", C.c_double), ("alg_gbps", C.c_double)]


def evaluate(ctx: Context, mat: CocktailMatrix, vec, res, coores, ntimes: int = 1000, verbose: bool = False):
    """evaluate (A/src/spmv_cocktail.cpp:1431-1455) on a context: upload,
    one checked do_spmv against coores, ntimes timed; returns the report."""
    fp = C.POINTER(C.c_float)
    v = np.ascontiguousarray(vec, np.float32)
    r = np.ascontiguousarray(res, np.float32)
    c = np.ascontiguousarray(coores, np.float32)
    rep = EvalReport()
    check(lib().clspmv_evaluate_report(ctx.h, C.byref(mat.s), v.ctypes.data_as(fp), r.ctypes.data_as(fp),
                                       c.ctypes.data_as(fp), ntimes, int(verbose), C.byref(rep)), "evaluate")
    return {k: getattr(rep, k) for k, _ in EvalReport._fields_}
