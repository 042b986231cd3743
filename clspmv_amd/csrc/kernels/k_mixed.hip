// Mixed-precision SpMV (y += A*x): matrix values stored in 4, 2 or 1 bytes,
// products and sums in FP32.  Value codecs (host encoders: host/codec.hpp):
//   kCodecBf16x2    v = hi + lo, two BF16 halves of one 32-bit word
//                   (hi = bf16(v), lo = bf16(v - hi)): same bytes as FP32,
//                   about 16 significant bits;
//   kCodecFp16Row   v = s_row * h, h FP16, one FP32 scale per row (per block
//                   row for BSR): 2 bytes per value;
//   kCodecFp8Slice  v = s * q, q FP8 E4M3 (OCP: bias 7, 3 mantissa bits,
//                   largest finite 448), one FP32 scale per 64-row slice of
//                   each diagonal (DIA) or of each block element (BSR):
//                   1 byte per value.
// DIA: one row per lane (a wave's value loads per diagonal stay contiguous:
// 256, 128 or 64 bytes).  BSR: the sliced layout of bsr_sliced_kernel
// with the encoded values and, optionally, block columns stored as int16
// deltas from the previous block of the same block row (first one from a
// per-block-row base).
#include <hip/hip_fp16.h>

#include "kcommon.hpp"
#include "launch.hpp"

namespace clspmv {
namespace dev {

__device__ __forceinline__ float bf16x2_value(uint32_t u) {
    return __uint_as_float(u & 0xffff0000u) + __uint_as_float(u << 16);
}
__device__ __forceinline__ float f16_value(uint16_t h) { return __half2float(__ushort_as_half(h)); }
// E4M3 decode from its definition: normal (e >= 1): 2^(e-7) * (1 + m/8),
// subnormal (e = 0): m * 2^-9; the caller's scale absorbs the magnitude.
__device__ __forceinline__ float e4m3_value(uint32_t q) {
    const uint32_t e = (q >> 3) & 15u, m = q & 7u;
    const float mag = e ? __uint_as_float(((e + 120u) << 23) | (m << 20)) : (float)m * (1.0f / 512.0f);
    return (q & 0x80u) ? -mag : mag;
}

__device__ __forceinline__ float xload_m(const float* __restrict__ x, int c, int cols) {
    return (c >= 0 && c < cols) ? x[c] : 0.f;
}

template <int CODEC>
__device__ __forceinline__ float dia_value(const void* __restrict__ data, const float* __restrict__ scale, size_t i,
                                          int d, int nsl, int row, float rs) {
    if (CODEC == kCodecBf16x2) return bf16x2_value(ld_stream(reinterpret_cast<const uint32_t*>(data) + i));
    if (CODEC == kCodecFp16Row) return rs * f16_value(reinterpret_cast<const uint16_t*>(data)[i]);
    // the 64 rows of a wave share one slice: a wave-uniform (scalar) load
    const int sl = __builtin_amdgcn_readfirstlane(row >> 6);
    return scale[(size_t)d * nsl + sl] * e4m3_value(reinterpret_cast<const uint8_t*>(data)[i]);
}

// One row per lane (as dia_row_kernel), diagonals four at a time so that
// four value loads and four x loads are in flight per lane.
template <int CODEC>
__global__ __launch_bounds__(256) void dia_mixed_kernel(const int* __restrict__ offsets, const void* __restrict__ data,
                                                        const float* __restrict__ scale, int L, int ndiag, int nsl,
                                                        const float* __restrict__ x, float* __restrict__ y, int rows,
                                                        int cols) {
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int row = tile * 256 + threadIdx.x;
    if (row >= rows) return;
    const float rs = CODEC == kCodecFp16Row ? scale[row] : 1.f;
    float acc = y[row];
    int d = 0;
    for (; d + 4 <= ndiag; d += 4) {
        float v[4], xv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            v[u] = dia_value<CODEC>(data, scale, (size_t)(d + u) * L + row, d + u, nsl, row, rs);
            xv[u] = xload_m(x, row + offsets[d + u], cols);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) acc = fmaf(v[u], xv[u], acc);
    }
    for (; d < ndiag; ++d)
        acc = fmaf(dia_value<CODEC>(data, scale, (size_t)d * L + row, d, nsl, row, rs), xload_m(x, row + offsets[d], cols),
                   acc);
    y[row] = acc;
}

template <int BH, int BW, int CODEC, bool DELTA>
__global__ __launch_bounds__(256) void bsr_mixed_kernel(const int* __restrict__ sp, const void* __restrict__ colv,
                                                        const int* __restrict__ base, const void* __restrict__ data,
                                                        const float* __restrict__ scale, int stride, int brows,
                                                        int nsl, const float* __restrict__ x, float* __restrict__ y,
                                                        int rows, int cols) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= brows) return;
    // a wave's 64 block rows are one slice: wave-uniform slice quantities
    const int s = __builtin_amdgcn_readfirstlane(r >> 6), t = r & 63;
    const int b = sp[s];
    const int width = (sp[s + 1] - b) >> 6;
    const float rs = CODEC == kCodecFp16Row ? scale[r] : 1.f;
    float acc[BH];
#pragma unroll
    for (int h = 0; h < BH; ++h) acc[h] = 0.f;
    int c = DELTA ? base[r] : 0;
#pragma unroll 3
    for (int k = 0; k < width; ++k) {
        const int idx = b + (k << 6) + t;
        if (DELTA)
            c += reinterpret_cast<const int16_t*>(colv)[idx];
        else
            c = reinterpret_cast<const int*>(colv)[idx];
        float xv[BW];
#pragma unroll
        for (int w = 0; w < BW; ++w) xv[w] = (c + w < cols) ? x[c + w] : 0.f;
#pragma unroll
        for (int h = 0; h < BH; ++h)
#pragma unroll
            for (int w = 0; w < BW; ++w) {
                const size_t e = (size_t)(h * BW + w) * stride + idx;
                float v;
                if (CODEC == kCodecBf16x2)
                    v = bf16x2_value(reinterpret_cast<const uint32_t*>(data)[e]);
                else if (CODEC == kCodecFp16Row)
                    v = rs * f16_value(reinterpret_cast<const uint16_t*>(data)[e]);
                else
                    v = scale[(size_t)(h * BW + w) * nsl + s] * e4m3_value(reinterpret_cast<const uint8_t*>(data)[e]);
                acc[h] = fmaf(v, xv[w], acc[h]);
            }
    }
#pragma unroll
    for (int h = 0; h < BH; ++h) {
        const int row = r * BH + h;
        if (row < rows) y[row] += acc[h];
    }
}

}  // namespace dev

int launch_dia_mixed(const DevDiaMixed& m, const float* x, float* y, hipStream_t st) {
    if (m.rows == 0) return 0;
    const int groups = (m.rows + 255) / 256;
    switch (m.codec) {
        case kCodecBf16x2:
            hipLaunchKernelGGL(dev::dia_mixed_kernel<kCodecBf16x2>, dim3(groups), dim3(256), 0, st, m.offsets, m.data,
                               m.scale, m.L, m.ndiag, m.nsl, x, y, m.rows, m.cols);
            break;
        case kCodecFp16Row:
            hipLaunchKernelGGL(dev::dia_mixed_kernel<kCodecFp16Row>, dim3(groups), dim3(256), 0, st, m.offsets,
                               m.data, m.scale, m.L, m.ndiag, m.nsl, x, y, m.rows, m.cols);
            break;
        default:
            hipLaunchKernelGGL(dev::dia_mixed_kernel<kCodecFp8Slice>, dim3(groups), dim3(256), 0, st, m.offsets,
                               m.data, m.scale, m.L, m.ndiag, m.nsl, x, y, m.rows, m.cols);
            break;
    }
    return (int)hipGetLastError();
}

template <int BH, int BW>
static void bsr_mixed_launch(const DevBsrMixed& m, const float* x, float* y, hipStream_t st) {
    const dim3 g((m.brows + 255) / 256), b(256);
#define CLSPMV_BSR_MIXED(C, D)                                                                                   \
    hipLaunchKernelGGL((dev::bsr_mixed_kernel<BH, BW, C, D>), g, b, 0, st, m.sp, m.col, m.base, m.data, m.scale, \
                       m.stride, m.brows, m.nsl, x, y, m.rows, m.cols)
    if (m.delta) {
        if (m.codec == kCodecBf16x2) CLSPMV_BSR_MIXED(kCodecBf16x2, true);
        else if (m.codec == kCodecFp16Row) CLSPMV_BSR_MIXED(kCodecFp16Row, true);
        else CLSPMV_BSR_MIXED(kCodecFp8Slice, true);
    } else {
        if (m.codec == kCodecBf16x2) CLSPMV_BSR_MIXED(kCodecBf16x2, false);
        else if (m.codec == kCodecFp16Row) CLSPMV_BSR_MIXED(kCodecFp16Row, false);
        else CLSPMV_BSR_MIXED(kCodecFp8Slice, false);
    }
#undef CLSPMV_BSR_MIXED
}

int launch_bsr_mixed(const DevBsrMixed& m, const float* x, float* y, hipStream_t st) {
    if (m.brows == 0) return 0;
    switch (m.bh * 10 + m.bw) {
        case 22: bsr_mixed_launch<2, 2>(m, x, y, st); break;
        case 33: bsr_mixed_launch<3, 3>(m, x, y, st); break;
        case 44: bsr_mixed_launch<4, 4>(m, x, y, st); break;
        default: return (int)hipErrorInvalidValue;
    }
    return (int)hipGetLastError();
}

}  // namespace clspmv
