"""ctypes binding of libclspmv.so (the C ABI declared in include/clspmv.h).

The library is loaded from clspmv_amd/lib/libclspmv.so (built in-tree by
``make -C clspmv_amd/csrc`` or ``__graft_entry__.build()``).  Loading fails
loudly when it is missing: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libclspmv.so")

OK = 0
VAL_BF16X2, VAL_FP16_ROW, VAL_FP8_SLICE, VAL_I16_ROW = 1, 2, 3, 4
E_INVALID, E_TOO_LARGE, E_PADDING, E_HIP, E_NOMEM, E_IO, E_SORT, E_UNSUPPORTED, E_NODEVICE = (
    -1, -2, -3, -4, -5, -6, -7, -8, -9)
ERROR_NAMES = {0: "OK", -1: "E_INVALID", -2: "E_TOO_LARGE", -3: "E_PADDING", -4: "E_HIP",
               -5: "E_NOMEM", -6: "E_IO", -7: "E_SORT", -8: "E_UNSUPPORTED", -9: "E_NODEVICE"}

FMT_CSR_SCALAR, FMT_CSR_VECTOR, FMT_BDIA, FMT_DIA, FMT_ELL, FMT_COO = 1, 2, 3, 4, 5, 6
FMT_BELL, FMT_BCSR, FMT_SELL, FMT_SBELL, FMT_BSR = 7, 8, 9, 10, 11

GPU_ALIGNMENT = 32
MAX_MEM_OBJ = 268435456

_IP = C.POINTER(C.c_int)
_FP = C.POINTER(C.c_float)


class MatInfo(C.Structure):
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("nnz", C.c_int)]


class Coo(C.Structure):
    _fields_ = [("matinfo", MatInfo), ("coo_row_id", _IP), ("coo_col_id", _IP), ("coo_data", _FP)]


class Csr(C.Structure):
    _fields_ = [("matinfo", MatInfo), ("csr_row_ptr", _IP), ("csr_col_id", _IP), ("csr_data", _FP)]


class Dia(C.Structure):
    _fields_ = [("matinfo", MatInfo), ("dia_num", C.c_int), ("dia_length", C.c_int),
                ("dia_length_aligned", C.c_int), ("dia_offsets", _IP), ("dia_data", _FP)]


class Bdia(C.Structure):
    _fields_ = [("matinfo", MatInfo), ("bdia_band_num", C.c_int), ("bdia_length", C.c_int),
                ("bdia_length_aligned", C.c_int), ("bdia_offsets", _IP), ("bdia_bptr", _IP),
                ("bdia_data", _FP)]


class Ell(C.Structure):
    _fields_ = [("matinfo", MatInfo), ("ell_num", C.c_int), ("ell_height_aligned", C.c_int),
                ("ell_col_id", _IP), ("ell_data", _FP)]


class Sell(C.Structure):
    _fields_ = [("matinfo", MatInfo), ("sell_slice_height", C.c_uint), ("sell_slice_num", C.c_int),
                ("sell_slice_ptr", _IP), ("sell_col_id", _IP), ("sell_data", _FP)]


class B4Csr(C.Structure):
    _fields_ = [("matinfo", MatInfo), ("b4csr_bwidth", C.c_uint), ("b4csr_bheight", C.c_uint),
                ("b4csr_row_num", C.c_int), ("b4csr_aligned_size", C.c_int), ("b4csr_block_num", C.c_int),
                ("b4csr_row_ptr", _IP), ("b4csr_col_id", _IP), ("b4csr_data", _FP)]


class B4Ell(C.Structure):
    _fields_ = [("matinfo", MatInfo), ("b4ell_bwidth", C.c_uint), ("b4ell_bheight", C.c_uint),
                ("b4ell_row_num", C.c_int), ("b4ell_block_num", C.c_int), ("b4ell_height_aligned", C.c_int),
                ("b4ell_float4_aligned", C.c_int), ("b4ell_col_id", _IP), ("b4ell_data", _FP)]


class Sbell(C.Structure):
    _fields_ = [("matinfo", MatInfo), ("sbell_bwidth", C.c_uint), ("sbell_bheight", C.c_uint),
                ("sbell_slice_height", C.c_uint), ("sbell_slice_num", C.c_int), ("sbell_row_num", C.c_int),
                ("sbell_slice_ptr", _IP), ("sbell_col_id", _IP), ("sbell_data", _FP)]


class Bsr(C.Structure):
    _fields_ = [("matinfo", MatInfo), ("bsr_bwidth", C.c_uint), ("bsr_bheight", C.c_uint),
                ("bsr_row_num", C.c_int), ("bsr_block_num", C.c_int), ("bsr_block_stride", C.c_int),
                ("bsr_row_ptr", _IP), ("bsr_col_id", _IP), ("bsr_data", _FP)]


class PlanPart(C.Structure):
    _fields_ = [("source", C.c_int), ("format", C.c_int), ("method", C.c_int), ("row0", C.c_int), ("row1", C.c_int),
                ("nnz", C.c_longlong), ("block", C.c_int)]


class XtileLayout(C.Structure):
    """clspmv_xtile_layout: host arrays of a CSR matrix's LDS x-tiled layout."""
    _fields_ = [("nblocks", C.c_int), ("groups", C.c_int), ("ntiles", C.c_int), ("nchunks", C.c_int),
                ("blk_row", C.POINTER(C.c_int)), ("blk_tile", C.POINTER(C.c_int)),
                ("tile_slice", C.POINTER(C.c_int)), ("tile_chunk", C.POINTER(C.c_int)),
                ("idx", C.POINTER(C.c_uint)), ("val", C.POINTER(C.c_float))]


class BatchEntry(C.Structure):
    _fields_ = [("csr", C.c_void_p), ("bsr", C.c_void_p), ("k", C.c_int)]


class SpmmTile(C.Structure):
    _fields_ = [("rows_cap", C.c_int), ("nnz_cap", C.c_int), ("long_threshold", C.c_int)]


class Cocktail(C.Structure):
    _fields_ = ([(f"ifuse{n}", C.c_ubyte) for n in ("bdia", "dia", "sbell", "bell", "bcsr", "sell", "ell", "csr", "coo")]
                + [(f"{n}_meth_num", C.c_int) for n in ("bdia", "dia", "sbell", "bell", "bcsr", "sell", "ell", "csr",
                                                        "coo")]
                + [("csr_group_num", C.c_int), ("coo_group_num", C.c_int), ("mat_width", C.c_int),
                   ("mat_height", C.c_int), ("mat_nnz", C.c_int), ("bdia", Bdia), ("dia", Dia), ("sbell", Sbell),
                   ("bell", B4Ell), ("bcsr", B4Csr), ("sell", Sell), ("ell", Ell), ("csr", Csr), ("coo", Coo)])


class MaxflopInfo(C.Structure):
    _fields_ = [("bdia_meth_num", C.c_int), ("bdia_max_flop", C.c_double), ("dia_meth_num", C.c_int),
                ("dia_max_flop", C.c_double), ("sbell_meth_num", C.c_int * 8), ("sbell_max_flop", C.c_double * 8),
                ("bell_meth_num", C.c_int * 8), ("bell_max_flop", C.c_double * 8), ("bcsr_meth_num", C.c_int * 8),
                ("bcsr_max_flop", C.c_double * 8), ("sell_meth_num", C.c_int), ("sell_max_flop", C.c_double),
                ("ell_meth_num", C.c_int), ("ell_max_flop", C.c_double), ("csr_meth_num", C.c_int),
                ("csr_groupnum", C.c_int), ("csr_max_flop", C.c_double), ("coo_meth_num", C.c_int),
                ("coo_groupnum", C.c_int), ("coo_max_flop", C.c_double), ("max_dia_base_flop", C.c_double),
                ("max_block_base_flop", C.c_double), ("max_flat_base_flop", C.c_double)]


class PartInfo(C.Structure):
    _fields_ = [("format", C.c_int), ("method", C.c_int), ("rows", C.c_int), ("cols", C.c_int),
                ("nnz", C.c_longlong), ("stored_bytes", C.c_longlong), ("alg_bytes", C.c_longlong),
                ("launches", C.c_int), ("kernel", C.c_char * 64), ("codec_bound", C.c_double)]


# (name, restype, argtypes)
_P = C.POINTER
_SIGS = [
    ("clspmv_last_error", C.c_char_p, []),
    ("clspmv_init_coo", None, [_P(Coo)]),
    ("clspmv_free_coo", None, [_P(Coo)]),
    ("clspmv_free_csr", None, [_P(Csr)]),
    ("clspmv_free_dia", None, [_P(Dia)]),
    ("clspmv_free_bdia", None, [_P(Bdia)]),
    ("clspmv_free_ell", None, [_P(Ell)]),
    ("clspmv_free_sell", None, [_P(Sell)]),
    ("clspmv_free_b4csr", None, [_P(B4Csr)]),
    ("clspmv_free_b4ell", None, [_P(B4Ell)]),
    ("clspmv_free_sbell", None, [_P(Sbell)]),
    ("clspmv_free_bsr", None, [_P(Bsr)]),
    ("clspmv_alloc_coo", C.c_int, [_P(Coo), C.c_int, C.c_int, C.c_int]),
    ("clspmv_if_sorted_coo", C.c_int, [_P(Coo)]),
    ("clspmv_sort_coo", C.c_int, [_P(Coo)]),
    ("clspmv_aligned_length", C.c_int, [C.c_int, C.c_int]),
    ("clspmv_coo2csr", C.c_int, [_P(Coo), _P(Csr)]),
    ("clspmv_coo2dia", C.c_int, [_P(Coo), _P(Dia), C.c_int]),
    ("clspmv_coo2bdia", C.c_int, [_P(Coo), _P(Bdia), C.c_int]),
    ("clspmv_coo2ell", C.c_int, [_P(Coo), _P(Ell), C.c_int, C.c_int]),
    ("clspmv_coo2sell", C.c_int, [_P(Coo), _P(Sell), C.c_uint]),
    ("clspmv_coo2b4csr", C.c_int, [_P(Coo), _P(B4Csr), C.c_uint, C.c_uint, C.c_int]),
    ("clspmv_coo2b4ell", C.c_int, [_P(Coo), _P(B4Ell), C.c_uint, C.c_uint, C.c_int, C.c_int]),
    ("clspmv_coo2sbell", C.c_int, [_P(Coo), _P(Sbell), C.c_uint, C.c_uint, C.c_uint]),
    ("clspmv_coo2bsr", C.c_int, [_P(Coo), _P(Bsr), C.c_uint, C.c_uint]),
    ("clspmv_pad_csr", C.c_int, [_P(Csr), _P(Csr), C.c_int]),
    ("clspmv_read_mmf", C.c_int, [C.c_char_p, _P(Coo)]),
    ("clspmv_write_mmf", C.c_int, [C.c_char_p, _P(Coo)]),
    ("clspmv_write_coo_bin", C.c_int, [C.c_char_p, _P(Coo)]),
    ("clspmv_read_coo_bin", C.c_int, [C.c_char_p, _P(Coo)]),
    ("clspmv_gen_laplacian2d", C.c_int, [_P(Coo), C.c_int, C.c_int]),
    ("clspmv_gen_stencil27", C.c_int, [_P(Coo), C.c_int, C.c_uint64]),
    ("clspmv_gen_uniform_rows", C.c_int, [_P(Coo), C.c_int, C.c_int, C.c_uint64]),
    ("clspmv_gen_zipf_rows", C.c_int, [_P(Coo), C.c_int, C.c_double, C.c_double, C.c_int, C.c_uint64]),
    ("clspmv_gen_block_rows", C.c_int, [_P(Coo), C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64]),
    ("clspmv_gen_cocktail_composite", C.c_int, [_P(Coo), C.c_int, C.c_uint64]),
    ("clspmv_gen_vector", C.c_int, [_FP, C.c_int, C.c_uint64]),
    ("clspmv_init", C.c_int, [C.c_int, _P(C.c_void_p)]),
    ("clspmv_finalize", C.c_int, [C.c_void_p]),
    ("clspmv_set_stream", C.c_int, [C.c_void_p, C.c_void_p]),
    ("clspmv_synchronize", C.c_int, [C.c_void_p]),
    ("clspmv_device_alloc", C.c_int, [C.c_void_p, C.c_size_t, _P(C.c_void_p)]),
    ("clspmv_device_free", C.c_int, [C.c_void_p, C.c_void_p]),
    ("clspmv_memcpy_h2d", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    ("clspmv_memcpy_d2h", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    ("clspmv_memcpy_d2d", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    ("clspmv_event_elapsed_begin", C.c_int, [C.c_void_p]),
    ("clspmv_event_elapsed_end", C.c_int, [C.c_void_p, _FP]),
    ("clspmv_part_csr_scalar", C.c_int, [C.c_void_p, _P(Csr), C.c_int, _P(C.c_void_p)]),
    ("clspmv_part_csr_vector", C.c_int, [C.c_void_p, _P(Csr), C.c_int, _P(C.c_void_p)]),
    ("clspmv_part_dia", C.c_int, [C.c_void_p, _P(Dia), C.c_int, _P(C.c_void_p)]),
    ("clspmv_part_bdia", C.c_int, [C.c_void_p, _P(Bdia), C.c_int, _P(C.c_void_p)]),
    ("clspmv_part_ell", C.c_int, [C.c_void_p, _P(Ell), C.c_int, _P(C.c_void_p)]),
    ("clspmv_part_sell", C.c_int, [C.c_void_p, _P(Sell), C.c_int, _P(C.c_void_p)]),
    ("clspmv_part_coo", C.c_int, [C.c_void_p, _P(Coo), C.c_int, _P(C.c_void_p)]),
    ("clspmv_part_bell", C.c_int, [C.c_void_p, _P(B4Ell), C.c_int, _P(C.c_void_p)]),
    ("clspmv_part_sbell", C.c_int, [C.c_void_p, _P(Sbell), C.c_int, _P(C.c_void_p)]),
    ("clspmv_part_bcsr", C.c_int, [C.c_void_p, _P(B4Csr), C.c_int, _P(C.c_void_p)]),
    ("clspmv_part_bsr", C.c_int, [C.c_void_p, _P(Bsr), C.c_int, _P(C.c_void_p)]),
    ("clspmv_part_spmv", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]),
    ("clspmv_set_option", C.c_int, [C.c_char_p, C.c_longlong]),
    ("clspmv_part_get_info", C.c_int, [C.c_void_p, _P(PartInfo)]),
    ("clspmv_part_free", C.c_int, [C.c_void_p]),
    ("clspmv_init_cocktail", None, [_P(Cocktail)]),
    ("clspmv_free_cocktail", None, [_P(Cocktail)]),
    ("clspmv_bench_load", C.c_int, [C.c_char_p, _P(C.c_void_p)]),
    ("clspmv_bench_free", C.c_int, [C.c_void_p]),
    ("clspmv_bench_sizes", C.c_int, [C.c_void_p, _P(C.c_int)]),
    ("clspmv_maxflop", C.c_int, [C.c_void_p, C.c_int, C.c_int, _P(MaxflopInfo)]),
    ("clspmv_find_overhead", C.c_double, [C.c_void_p, C.c_int, C.c_int]),
    ("clspmv_analyze_matrix", C.c_int, [C.c_void_p, _P(Coo), _P(Cocktail), C.c_float, C.c_int, C.c_int, C.c_int]),
    ("clspmv_init_mat_kernels", C.c_int, [C.c_void_p, _P(Cocktail), _FP, _FP, C.c_int, _P(C.c_void_p)]),
    ("clspmv_free_mat_kernels", C.c_int, [C.c_void_p]),
    ("clspmv_do_spmv", C.c_int, [C.c_void_p, C.c_int]),
    ("clspmv_analyze_plan", C.c_int, [C.c_void_p, _P(Coo), C.c_float, C.c_int, C.c_int, _P(C.c_void_p)]),
    ("clspmv_plan_cocktail", C.c_int, [C.c_void_p, _P(C.c_void_p)]),
    ("clspmv_plan_num_parts", C.c_int, [C.c_void_p]),
    ("clspmv_plan_get_part", C.c_int, [C.c_void_p, C.c_int, _P(PlanPart)]),
    ("clspmv_xtile_layout_csr", C.c_int, [_P(Csr), _P(XtileLayout)]),
    ("clspmv_free_xtile_layout", None, [_P(XtileLayout)]),
    ("clspmv_plan_partition", C.c_void_p, [C.c_void_p]),
    ("clspmv_plan_free", C.c_int, [C.c_void_p]),
    ("clspmv_init_plan_kernels", C.c_int, [C.c_void_p, C.c_void_p, _FP, _FP, _P(C.c_void_p)]),
    ("clspmv_power_iterate", C.c_int, [C.c_void_p, C.c_int]),
    ("clspmv_evaluate", C.c_int, [C.c_void_p, _FP, _FP, _FP]),
    ("clspmv_evaluate_report", C.c_int, [C.c_void_p, C.c_void_p, _FP, _FP, _FP, C.c_int, C.c_int, C.c_void_p]),
    ("clspmv_cpy_vector_from_cpu", C.c_int, [C.c_void_p, _FP, C.c_int]),
    ("clspmv_cpy_vector_from_gpu", C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    ("clspmv_cpy_result_from_cpu", C.c_int, [C.c_void_p, _FP, C.c_int]),
    ("clspmv_cpy_result_from_gpu", C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    ("clspmv_cpy_result_to_cpu", C.c_int, [C.c_void_p, _FP, C.c_int]),
    ("clspmv_cpy_result_to_gpu", C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    ("clspmv_gpumat_buffers", C.c_int, [C.c_void_p, _P(C.c_void_p), _P(C.c_void_p)]),
    ("clspmv_gpumat_num_parts", C.c_int, [C.c_void_p]),
    ("clspmv_gpumat_part_info", C.c_int, [C.c_void_p, C.c_int, _P(PartInfo)]),
    ("clspmv_part_dia_mixed", C.c_int, [C.c_void_p, _P(Dia), C.c_int, _P(C.c_void_p)]),
    ("clspmv_part_bsr_mixed", C.c_int, [C.c_void_p, _P(Bsr), C.c_int, C.c_int, _P(C.c_void_p)]),
    ("clspmv_part_csr_mixed", C.c_int, [C.c_void_p, _P(Csr), C.c_int, _P(C.c_void_p)]),
    ("clspmv_spmm_default_tile", SpmmTile, [C.c_int]),
    ("clspmv_batch_create", C.c_int, [C.c_void_p, C.c_int, _P(_P(Csr)), _P(C.c_int), _P(SpmmTile), _P(C.c_void_p)]),
    ("clspmv_batch_create_entries", C.c_int, [C.c_void_p, C.c_int, C.c_void_p, _P(SpmmTile), _P(C.c_void_p)]),
    ("clspmv_batch_free", C.c_int, [C.c_void_p]),
    ("clspmv_batch_size", C.c_int, [C.c_void_p]),
    ("clspmv_batch_buffers", C.c_int, [C.c_void_p, C.c_int, _P(C.c_void_p), _P(C.c_void_p)]),
    ("clspmv_batch_set_x", C.c_int, [C.c_void_p, C.c_int, _FP]),
    ("clspmv_batch_set_y", C.c_int, [C.c_void_p, C.c_int, _FP]),
    ("clspmv_batch_get_y", C.c_int, [C.c_void_p, C.c_int, _FP]),
    ("clspmv_batch_run", C.c_int, [C.c_void_p, C.c_int]),
    ("clspmv_batch_capture", C.c_int, [C.c_void_p]),
    ("clspmv_batch_stats", C.c_int, [C.c_void_p, _P(C.c_longlong), _P(C.c_longlong), _P(C.c_longlong),
                                     _P(C.c_int), _P(C.c_int)]),
    ("clspmv_mlp_create", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_uint64, _P(C.c_void_p)]),
    ("clspmv_mlp_free", C.c_int, [C.c_void_p]),
    ("clspmv_mlp_num_params", C.c_longlong, [C.c_void_p]),
    ("clspmv_mlp_get_params", C.c_int, [C.c_void_p, _FP]),
    ("clspmv_mlp_set_params", C.c_int, [C.c_void_p, _FP]),
    ("clspmv_mlp_loss_grad", C.c_int, [C.c_void_p, _FP, _IP, C.c_int, _FP, _FP]),
    ("clspmv_mlp_train", C.c_int, [C.c_void_p, _FP, _IP, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float,
                                   C.c_float, _FP]),
    ("clspmv_mlp_predict", C.c_int, [C.c_void_p, _FP, C.c_int, _FP, _IP]),
    ("clspmv_probe_copy", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, _FP]),
    ("clspmv_probe_launch", C.c_int, [C.c_void_p, C.c_int, _FP]),
    ("clspmv_probe_launch_grid", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, _FP]),
    ("clspmv_probe_graph_launch", C.c_int, [C.c_void_p, C.c_int, C.c_int, _FP]),
    ("clspmv_part_xtile_stats", C.c_int, [C.c_void_p, _P(C.c_longlong), _P(C.c_longlong)]),
    ("clspmv_coo_spmv_noalloc", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                          C.c_void_p, C.c_void_p]),
]

_lib = None


def lib() -> C.CDLL:
    """Load libclspmv.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `make -C clspmv_amd/csrc` "
                "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, res, args in _SIGS:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    """Names this binding declares (all must be exported by the library)."""
    return [s[0] for s in _SIGS]


class ClspmvError(RuntimeError):
    def __init__(self, code: int, what: str):
        msg = lib().clspmv_last_error()
        detail = msg.decode() if msg else ""
        super().__init__(f"{what} failed: {ERROR_NAMES.get(code, code)} {detail}")
        self.code = code


def check(code: int, what: str) -> int:
    if code != OK:
        raise ClspmvError(code, what)
    return code
