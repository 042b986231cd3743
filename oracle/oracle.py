"""Python loader of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker.  ``liboracle.so`` is the plain-C
restatement of the reference's CPU SpMV (spmv_oracle.c); ``ref()`` returns
the reference's own compiled CPU code (oracle/_ref/libclspmv_ref.so) when it
was built in this container, else None.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libclspmv_ref.so")

_IP = C.POINTER(C.c_int)
_FP = C.POINTER(C.c_float)
_DP = C.POINTER(C.c_double)

_o = None
_r = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _o
    if _o is None:
        if not os.path.exists(ORACLE_SO):
            build()
        L = C.CDLL(ORACLE_SO)
        sigs = {
            "oracle_spmv_only": [C.c_int, C.c_int, _IP, _IP, _FP, _FP, _FP],
            "oracle_coo_spmv": [C.c_int, _IP, _IP, _FP, _FP, _FP],
            "oracle_coo_spmv_f64": [C.c_int, C.c_int, _IP, _IP, _FP, _FP, _DP],
            "oracle_abs_row_sum": [C.c_int, C.c_int, _IP, _IP, _FP, _FP, _DP],
            "oracle_dia_spmv": [C.c_int, C.c_int, _IP, _FP, _FP, C.c_int, C.c_int, _FP],
            "oracle_bdia_spmv": [C.c_int, C.c_int, _IP, _IP, _FP, _FP, C.c_int, C.c_int, _FP],
            "oracle_ell_spmv": [C.c_int, C.c_int, _IP, _FP, _FP, C.c_int, _FP],
            "oracle_sell_spmv": [C.c_int, C.c_int, _IP, _IP, _FP, _FP, C.c_int, _FP],
            "oracle_csr_spmv": [C.c_int, _IP, _IP, _FP, _FP, _FP],
            "oracle_b4csr_spmv": [C.c_int, C.c_int, C.c_int, C.c_int, _IP, _IP, _FP, _FP, C.c_int, C.c_int, _FP],
            "oracle_b4ell_spmv": [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _IP, _FP, _FP, C.c_int,
                                  C.c_int, _FP],
            "oracle_sbell_spmv": [C.c_int, C.c_int, C.c_int, C.c_int, _IP, _IP, _FP, _FP, C.c_int, C.c_int, _FP],
            "oracle_bsr_spmv": [C.c_int, C.c_int, C.c_int, C.c_int, _IP, _IP, _FP, _FP, C.c_int, C.c_int, _FP],
            "oracle_two_vec_compare": [_FP, _FP, C.c_int, _DP],
        }
        for name, args in sigs.items():
            f = getattr(L, name)
            f.restype = None
            f.argtypes = args
        L.oracle_csr_spmv_omp.restype = C.c_int
        L.oracle_csr_spmv_omp.argtypes = [C.c_int, _IP, _IP, _FP, _FP, _FP]
        _o = L
    return _o


def ref():
    """The reference's compiled CPU code, or None when it was not built."""
    global _r
    if _r is None and os.path.exists(REF_SO):
        R = C.CDLL(REF_SO)
        vp = C.c_void_p
        for name, res, args in [
            ("ref_sort_coo", C.c_int, [vp]), ("ref_coo2csr", C.c_int, [vp, vp]),
            ("ref_coo2dia", C.c_int, [vp, vp, C.c_int]), ("ref_coo2bdia", C.c_int, [vp, vp, C.c_int]),
            ("ref_coo2ell", C.c_int, [vp, vp, C.c_int, C.c_int]), ("ref_coo2sell", C.c_int, [vp, vp, C.c_uint]),
            ("ref_coo2b4csr", C.c_int, [vp, vp, C.c_uint, C.c_uint, C.c_int]),
            ("ref_coo2b4ell", C.c_int, [vp, vp, C.c_uint, C.c_uint, C.c_int, C.c_int]),
            ("ref_coo2sbell", C.c_int, [vp, vp, C.c_uint, C.c_uint, C.c_uint]),
            ("ref_pad_csr", None, [vp, vp, C.c_int]), ("ref_read_mmf", None, [C.c_char_p, vp]),
            ("ref_spmv_only", None, [vp, _FP, _FP]), ("ref_free", None, [vp]),
        ]:
            f = getattr(R, name)
            f.restype = res
            f.argtypes = args
        _r = R
    return _r


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def spmv_only(rows, cols, vals, x, height):
    """The reference's CPU SpMV restated: FP32 y = A x in COO order."""
    rows = np.ascontiguousarray(rows, np.int32)
    cols = np.ascontiguousarray(cols, np.int32)
    vals = np.ascontiguousarray(vals, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    y = np.empty(height, np.float32)
    lib().oracle_spmv_only(rows.size, height, _p(rows, C.c_int), _p(cols, C.c_int), _p(vals, C.c_float),
                           _p(x, C.c_float), _p(y, C.c_float))
    return y


def spmv_f64(rows, cols, vals, x, height):
    rows = np.ascontiguousarray(rows, np.int32)
    cols = np.ascontiguousarray(cols, np.int32)
    vals = np.ascontiguousarray(vals, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    y = np.empty(height, np.float64)
    lib().oracle_coo_spmv_f64(rows.size, height, _p(rows, C.c_int), _p(cols, C.c_int), _p(vals, C.c_float),
                              _p(x, C.c_float), _p(y, C.c_double))
    return y


def abs_row_sum(rows, cols, vals, x, height):
    rows = np.ascontiguousarray(rows, np.int32)
    cols = np.ascontiguousarray(cols, np.int32)
    vals = np.ascontiguousarray(vals, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    s = np.empty(height, np.float64)
    lib().oracle_abs_row_sum(rows.size, height, _p(rows, C.c_int), _p(cols, C.c_int), _p(vals, C.c_float),
                             _p(x, C.c_float), _p(s, C.c_double))
    return s


def csr_spmv_omp(rp, ci, cv, x):
    """Project CPU baseline (not the reference): FP32 CSR SpMV on all host
    cores with OpenMP.  Returns (y, threads used)."""
    rp = np.ascontiguousarray(rp, np.int32)
    ci = np.ascontiguousarray(ci, np.int32)
    cv = np.ascontiguousarray(cv, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    y = np.empty(rp.size - 1, np.float32)
    n = lib().oracle_csr_spmv_omp(rp.size - 1, _p(rp, C.c_int), _p(ci, C.c_int), _p(cv, C.c_float),
                                  _p(x, C.c_float), _p(y, C.c_float))
    return y, n


def two_vec_compare(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    out = np.zeros(3, np.float64)
    lib().oracle_two_vec_compare(_p(a, C.c_float), _p(b, C.c_float), a.size, _p(out, C.c_double))
    return {"distance": out[0], "max_diff": out[1], "max_ratio": out[2]}


RTOL = 1e-5  # FP32 parity bound against the reference CPU SpMV (per element)


def parity_report(y_gpu, coo, x, y0=None):
    """Compare a GPU y (= y0 + A x) with the oracle.

    Per element, with y_ref the reference-order FP32 result, y64 the FP64
    oracle and S = sum_j |a_ij x_j| (the scale of the row's rounding error):
        err_ref = |y_gpu - y_ref| / max(|y_ref|, S, tiny)
        err_64  = |y_gpu - y64|   / max(|y64|,  S, tiny)
    An element passes when err_ref <= 1e-5, or when the GPU value is at least
    as close to the FP64 value as the reference-order FP32 value is (long
    rows, where the sequential FP32 sum itself drifts from the exact sum).
    """
    rows, cols, vals = coo.row_id, coo.col_id, coo.data
    h = coo.height
    y_ref = spmv_only(rows, cols, vals, x, h)
    y64 = spmv_f64(rows, cols, vals, x, h)
    S = abs_row_sum(rows, cols, vals, x, h)
    if y0 is not None:
        y_ref = (y_ref.astype(np.float32) + np.asarray(y0, np.float32)).astype(np.float32)
        y64 = y64 + np.asarray(y0, np.float64)
        S = S + np.abs(np.asarray(y0, np.float64))
    yg = np.asarray(y_gpu, np.float64)
    tiny = 1e-30
    err_ref = np.abs(yg - y_ref) / np.maximum(np.maximum(np.abs(y_ref), S), tiny)
    err_64 = np.abs(yg - y64) / np.maximum(np.maximum(np.abs(y64), S), tiny)
    ref_err_64 = np.abs(y_ref.astype(np.float64) - y64) / np.maximum(np.maximum(np.abs(y64), S), tiny)
    ok = (err_ref <= RTOL) | (err_64 <= ref_err_64 + RTOL * 0.1)
    return {
        "max_err_ref": float(err_ref.max(initial=0.0)),
        "max_err_f64": float(err_64.max(initial=0.0)),
        "max_ref_err_f64": float(ref_err_64.max(initial=0.0)),
        "n_fail": int((~ok).sum()),
        "ok": bool(ok.all()),
    }
