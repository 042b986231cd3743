"""Host-side mirror of the reference's sparse data model and GPU parts.

Python names follow the reference (A/include/matrix_storage.h, spmv_all
drivers): ``coo2csr``, ``coo2dia``, ``coo2ell`` ... return objects that own
the C structs filled by libclspmv; ``Context`` and ``Part`` wrap the HIP
runtime (device buffers, one part per uploaded format).  All compute runs in
libclspmv (C++ conversions, HIP kernels); nothing here computes an SpMV.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _lib as L
from ._lib import check, lib


def _view(ptr, n, dtype):
    """numpy view of a C array (no copy); empty when n == 0."""
    if n <= 0 or not ptr:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(int(n),))


class _Owned:
    """Owns one C struct whose arrays were malloc'd by libclspmv."""

    _free = None

    def __init__(self, struct):
        self.s = struct
        self._alive = True

    def free(self):
        if getattr(self, "_alive", False):
            getattr(lib(), self._free)(C.byref(self.s))
            self._alive = False

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    @property
    def width(self):
        return self.s.matinfo.width

    @property
    def height(self):
        return self.s.matinfo.height

    @property
    def nnz(self):
        return self.s.matinfo.nnz


class CooMatrix(_Owned):
    _free = "clspmv_free_coo"

    def __init__(self, struct=None):
        super().__init__(struct if struct is not None else L.Coo())

    @classmethod
    def from_arrays(cls, rows, cols, vals, width, height):
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        cols = np.ascontiguousarray(cols, dtype=np.int32)
        vals = np.ascontiguousarray(vals, dtype=np.float32)
        m = cls()
        check(lib().clspmv_alloc_coo(C.byref(m.s), int(width), int(height), int(rows.size)), "alloc_coo")
        m.row_id[:] = rows
        m.col_id[:] = cols
        m.data[:] = vals
        return m

    @classmethod
    def read_mmf(cls, path):
        m = cls()
        check(lib().clspmv_read_mmf(str(path).encode(), C.byref(m.s)), "ReadMMF")
        return m

    def write_mmf(self, path):
        check(lib().clspmv_write_mmf(str(path).encode(), C.byref(self.s)), "write_mmf")

    @property
    def row_id(self):
        return _view(self.s.coo_row_id, self.nnz, np.int32)

    @property
    def col_id(self):
        return _view(self.s.coo_col_id, self.nnz, np.int32)

    @property
    def data(self):
        return _view(self.s.coo_data, self.nnz, np.float32)

    def is_sorted(self):
        return bool(lib().clspmv_if_sorted_coo(C.byref(self.s)))

    def sort(self):
        check(lib().clspmv_sort_coo(C.byref(self.s)), "sort_coo")

    def copy(self):
        return CooMatrix.from_arrays(self.row_id, self.col_id, self.data, self.width, self.height)


def _gen(fn, *args):
    m = CooMatrix()
    check(getattr(lib(), fn)(C.byref(m.s), *args), fn)
    return m


def gen_laplacian2d(nx, ny):
    return _gen("clspmv_gen_laplacian2d", nx, ny)


def gen_stencil27(n, seed):
    return _gen("clspmv_gen_stencil27", n, C.c_uint64(seed))


def gen_uniform_rows(rows, per_row, seed):
    return _gen("clspmv_gen_uniform_rows", rows, per_row, C.c_uint64(seed))


def gen_zipf_rows(rows, exponent, mean, max_len, seed):
    return _gen("clspmv_gen_zipf_rows", rows, C.c_double(exponent), C.c_double(mean), max_len, C.c_uint64(seed))


def gen_block_rows(block_rows, bdim, blocks_per_row, band, seed):
    return _gen("clspmv_gen_block_rows", block_rows, bdim, blocks_per_row, band, C.c_uint64(seed))


def gen_cocktail_composite(rows, seed):
    return _gen("clspmv_gen_cocktail_composite", rows, C.c_uint64(seed))


def gen_vector(n, seed):
    x = np.empty(n, dtype=np.float32)
    check(lib().clspmv_gen_vector(x.ctypes.data_as(C.POINTER(C.c_float)), n, C.c_uint64(seed)), "gen_vector")
    return x


# ------------------------------------------------------------------ formats

class CsrMatrix(_Owned):
    _free = "clspmv_free_csr"

    @property
    def row_ptr(self):
        return _view(self.s.csr_row_ptr, self.height + 1, np.int32)

    @property
    def col_id(self):
        return _view(self.s.csr_col_id, self.nnz, np.int32)

    @property
    def data(self):
        return _view(self.s.csr_data, self.nnz, np.float32)


class DiaMatrix(_Owned):
    _free = "clspmv_free_dia"

    @property
    def offsets(self):
        return _view(self.s.dia_offsets, self.s.dia_num, np.int32)

    @property
    def data(self):
        return _view(self.s.dia_data, self.s.dia_num * self.s.dia_length_aligned, np.float32)


class BdiaMatrix(_Owned):
    _free = "clspmv_free_bdia"

    @property
    def offsets(self):
        return _view(self.s.bdia_offsets, self.s.bdia_band_num, np.int32)

    @property
    def bptr(self):
        return _view(self.s.bdia_bptr, self.s.bdia_band_num + 1, np.int32)

    @property
    def data(self):
        nd = int(self.bptr[-1]) if self.s.bdia_band_num else 0
        return _view(self.s.bdia_data, nd * self.s.bdia_length_aligned, np.float32)


class EllMatrix(_Owned):
    _free = "clspmv_free_ell"

    @property
    def col_id(self):
        return _view(self.s.ell_col_id, self.s.ell_num * self.s.ell_height_aligned, np.int32)

    @property
    def data(self):
        return _view(self.s.ell_data, self.s.ell_num * self.s.ell_height_aligned, np.float32)


class SellMatrix(_Owned):
    _free = "clspmv_free_sell"

    @property
    def slice_ptr(self):
        return _view(self.s.sell_slice_ptr, self.s.sell_slice_num + 1, np.int32)

    @property
    def col_id(self):
        return _view(self.s.sell_col_id, int(self.slice_ptr[-1]), np.int32)

    @property
    def data(self):
        return _view(self.s.sell_data, int(self.slice_ptr[-1]), np.float32)


class B4CsrMatrix(_Owned):
    _free = "clspmv_free_b4csr"

    @property
    def row_ptr(self):
        return _view(self.s.b4csr_row_ptr, self.s.b4csr_row_num + 1, np.int32)

    @property
    def col_id(self):
        return _view(self.s.b4csr_col_id, self.s.b4csr_block_num, np.int32)

    @property
    def data(self):
        n = self.s.b4csr_aligned_size * self.s.b4csr_bheight * (self.s.b4csr_bwidth // 4)
        return _view(self.s.b4csr_data, n, np.float32)


class B4EllMatrix(_Owned):
    _free = "clspmv_free_b4ell"

    @property
    def col_id(self):
        return _view(self.s.b4ell_col_id, self.s.b4ell_height_aligned * self.s.b4ell_block_num, np.int32)

    @property
    def data(self):
        s = self.s
        n = s.b4ell_float4_aligned * s.b4ell_bheight * (s.b4ell_bwidth // 4) * s.b4ell_block_num
        return _view(s.b4ell_data, n, np.float32)


class SbellMatrix(_Owned):
    _free = "clspmv_free_sbell"

    @property
    def slice_ptr(self):
        return _view(self.s.sbell_slice_ptr, self.s.sbell_slice_num + 1, np.int32)

    @property
    def col_id(self):
        return _view(self.s.sbell_col_id, int(self.slice_ptr[-1]), np.int32)

    @property
    def data(self):
        return _view(self.s.sbell_data, int(self.slice_ptr[-1]) * self.s.sbell_bwidth * self.s.sbell_bheight,
                     np.float32)


class BsrMatrix(_Owned):
    _free = "clspmv_free_bsr"

    @property
    def row_ptr(self):
        return _view(self.s.bsr_row_ptr, self.s.bsr_row_num + 1, np.int32)

    @property
    def col_id(self):
        return _view(self.s.bsr_col_id, self.s.bsr_block_num, np.int32)

    @property
    def data(self):
        return _view(self.s.bsr_data, self.s.bsr_block_stride * self.s.bsr_bwidth * self.s.bsr_bheight, np.float32)


def _convert(cls, struct_t, fn, coo, *args):
    out = cls(struct_t())
    rc = getattr(lib(), fn)(C.byref(coo.s), C.byref(out.s), *args)
    if rc != L.OK:
        out._alive = True
        out.free()
        raise L.ClspmvError(rc, fn)
    return out


def coo2csr(coo):
    return _convert(CsrMatrix, L.Csr, "clspmv_coo2csr", coo)


def coo2dia(coo, alignment=L.GPU_ALIGNMENT):
    return _convert(DiaMatrix, L.Dia, "clspmv_coo2dia", coo, alignment)


def coo2bdia(coo, alignment=L.GPU_ALIGNMENT):
    return _convert(BdiaMatrix, L.Bdia, "clspmv_coo2bdia", coo, alignment)


def coo2ell(coo, alignment=L.GPU_ALIGNMENT, ellnum=0):
    return _convert(EllMatrix, L.Ell, "clspmv_coo2ell", coo, alignment, ellnum)


def coo2sell(coo, slice_height=L.GPU_ALIGNMENT):
    return _convert(SellMatrix, L.Sell, "clspmv_coo2sell", coo, slice_height)


def coo2b4csr(coo, bwidth, bheight, alignment=L.GPU_ALIGNMENT):
    return _convert(B4CsrMatrix, L.B4Csr, "clspmv_coo2b4csr", coo, bwidth, bheight, alignment)


def coo2b4ell(coo, bwidth, bheight, alignment=L.GPU_ALIGNMENT, ellnum=0):
    return _convert(B4EllMatrix, L.B4Ell, "clspmv_coo2b4ell", coo, bwidth, bheight, alignment, ellnum)


def coo2sbell(coo, bwidth, bheight, slice_height=32):
    return _convert(SbellMatrix, L.Sbell, "clspmv_coo2sbell", coo, bwidth, bheight, slice_height)


def coo2bsr(coo, bwidth, bheight):
    return _convert(BsrMatrix, L.Bsr, "clspmv_coo2bsr", coo, bwidth, bheight)


def pad_csr(csr, alignment):
    out = CsrMatrix(L.Csr())
    check(lib().clspmv_pad_csr(C.byref(csr.s), C.byref(out.s), alignment), "pad_csr")
    return out


# ------------------------------------------------------------------ device

class Context:
    """HIP device + stream (replaces the reference's OpenCL context/queue)."""

    def __init__(self, device: int = 0):
        self.h = C.c_void_p()
        check(lib().clspmv_init(device, C.byref(self.h)), "clspmv_init")

    def close(self):
        if self.h:
            lib().clspmv_finalize(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        check(lib().clspmv_synchronize(self.h), "synchronize")

    def time_begin(self):
        check(lib().clspmv_event_elapsed_begin(self.h), "event begin")

    def time_end(self) -> float:
        ms = C.c_float()
        check(lib().clspmv_event_elapsed_end(self.h, C.byref(ms)), "event end")
        return float(ms.value)


class DeviceArray:
    """A device buffer managed through the C ABI (float32 or int32)."""

    def __init__(self, ctx: Context, n: int, dtype=np.float32):
        self.ctx, self.n, self.dtype = ctx, int(n), np.dtype(dtype)
        self.ptr = C.c_void_p()
        check(lib().clspmv_device_alloc(ctx.h, max(16, self.n * self.dtype.itemsize), C.byref(self.ptr)),
              "device_alloc")

    @classmethod
    def from_numpy(cls, ctx, a):
        a = np.ascontiguousarray(a)
        d = cls(ctx, a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        if self.n:
            check(lib().clspmv_memcpy_h2d(self.ctx.h, self.ptr, a.ctypes.data, a.nbytes), "h2d")

    def download(self):
        out = np.empty(self.n, dtype=self.dtype)
        if self.n:
            check(lib().clspmv_memcpy_d2h(self.ctx.h, out.ctypes.data, self.ptr, out.nbytes), "d2h")
        return out

    def free(self):
        # after Context.close() the device memory is gone with the context
        if self.ptr and self.ctx.h:
            lib().clspmv_device_free(self.ctx.h, self.ptr)
        self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


_PART_FN = {
    L.FMT_CSR_SCALAR: "clspmv_part_csr_scalar", L.FMT_CSR_VECTOR: "clspmv_part_csr_vector",
    L.FMT_DIA: "clspmv_part_dia", L.FMT_BDIA: "clspmv_part_bdia", L.FMT_ELL: "clspmv_part_ell",
    L.FMT_SELL: "clspmv_part_sell", L.FMT_COO: "clspmv_part_coo", L.FMT_BELL: "clspmv_part_bell",
    L.FMT_SBELL: "clspmv_part_sbell", L.FMT_BCSR: "clspmv_part_bcsr", L.FMT_BSR: "clspmv_part_bsr",
}


class Part:
    """One format uploaded to the device with its kernel variant."""

    def __init__(self, ctx: Context, fmt: int, mat, method: int = 0):
        self.ctx = ctx
        self.h = C.c_void_p()
        check(getattr(lib(), _PART_FN[fmt])(ctx.h, C.byref(mat.s), method, C.byref(self.h)), _PART_FN[fmt])
        inf = L.PartInfo()
        check(lib().clspmv_part_get_info(self.h, C.byref(inf)), "part_get_info")
        self.info = {k: getattr(inf, k) for k, _ in L.PartInfo._fields_}
        self.info["kernel"] = inf.kernel.decode()

    def spmv(self, x: DeviceArray, y: DeviceArray, ntimes: int = 1):
        """Enqueue ntimes * (y += A x) on the context stream."""
        check(lib().clspmv_part_spmv(self.h, x.ptr, y.ptr, ntimes), "part_spmv")

    def xtile_stats(self):
        """LDS x-tiled parts: tiles and staged x bytes per SpMV (0 otherwise)."""
        t, b = C.c_longlong(), C.c_longlong()
        check(lib().clspmv_part_xtile_stats(self.h, C.byref(t), C.byref(b)), "part_xtile_stats")
        return {"tiles": t.value, "staged_x_bytes": b.value}

    def free(self):
        if self.h and self.ctx.h:
            lib().clspmv_part_free(self.h)
        self.h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class MixedPart(Part):
    """A mixed-precision part (clspmv_part_dia_mixed / clspmv_part_bsr_mixed /
    clspmv_part_csr_mixed): values stored with `codec` (L.VAL_*), FP32
    accumulation.  CSR parts use the LDS x-tiled layout."""

    def __init__(self, ctx: Context, fmt: int, mat, codec: int, delta_cols: bool = True):
        self.ctx = ctx
        self.h = C.c_void_p()
        if fmt == L.FMT_DIA:
            check(lib().clspmv_part_dia_mixed(ctx.h, C.byref(mat.s), codec, C.byref(self.h)), "part_dia_mixed")
        elif fmt == L.FMT_BSR:
            check(lib().clspmv_part_bsr_mixed(ctx.h, C.byref(mat.s), codec, int(delta_cols), C.byref(self.h)),
                  "part_bsr_mixed")
        elif fmt in (L.FMT_CSR_VECTOR, L.FMT_CSR_SCALAR):
            check(lib().clspmv_part_csr_mixed(ctx.h, C.byref(mat.s), codec, C.byref(self.h)), "part_csr_mixed")
        else:
            raise ValueError("mixed precision is provided for DIA, BSR and CSR parts")
        inf = L.PartInfo()
        check(lib().clspmv_part_get_info(self.h, C.byref(inf)), "part_get_info")
        self.info = {k: getattr(inf, k) for k, _ in L.PartInfo._fields_}
        self.info["kernel"] = inf.kernel.decode()
