// Mixed-precision SpMV (y += A*x): matrix values stored in 4, 2 or 1 bytes,
// products and sums in FP32.  Value codecs (host encoders: host/codec.hpp):
//   kCodecBf16x2    v = hi + lo, two BF16 halves of one 32-bit word
//                   (hi = bf16(v), lo = bf16(v - hi)): same bytes as FP32,
//                   about 16 significant bits;
//   kCodecFp16Row   v = s_row * h, h FP16, one power-of-two FP32 scale per
//                   row (per block row for BSR), applied to the row sums:
//                   2 bytes per value;
//   kCodecFp8Slice  v = s * q, q FP8 E4M3 (OCP: bias 7, 3 mantissa bits,
//                   largest finite 448), one FP32 scale per slice: 256 rows
//                   of a diagonal (DIA) or 64 block rows of a block element
//                   (BSR): 1 byte per value;
//   kCodecI16Row    v = s_row * q, q a signed 16-bit integer (|q| <= 32767),
//                   s_row = (largest magnitude of the row) / 32767, applied
//                   to the row sums: 2 bytes per value, every value within
//                   s_row / 2 of its FP32 value.
// DIA: 1, 2 or 4 consecutive rows per lane, so that every lane loads one
// 32-bit word of values per diagonal.  BSR: the sliced layout of
// bsr_sliced_kernel with the encoded values and, optionally, block columns
// stored as int16 deltas from the previous block of the same block row
// (first one from a per-block-row base).
#include "kcodec.hpp"
#include "kcommon.hpp"
#include "launch.hpp"

namespace clspmv {
namespace dev {

// Loads of the 1- and 2-byte streams (FP16 / FP8 values, int16 columns):
// plain loads -- with the nontemporal hint these streams ran 1.3-1.4x
// slower on MI355X (c1a DIA FP16 35.9 vs 26.9 us, c1d BSR FP16 111 vs 81).
constexpr bool kNarrowNontemporal = false;
template <class T>
__device__ __forceinline__ T ld_narrow(const T* p) {
    if constexpr (kNarrowNontemporal) return ld_stream(p);
    return *p;
}

__device__ __forceinline__ float xload_m(const float* __restrict__ x, int c, int cols) {
    return (c >= 0 && c < cols) ? x[c] : 0.f;
}

// R consecutive floats at p (p aligned to 4 R bytes, 16 for R = 8) as one
// load / store (two for R = 8).
template <int R>
__device__ __forceinline__ void load_vec(const float* p, float (&v)[R]) {
    if constexpr (R == 8) {
        const float4 t = *reinterpret_cast<const float4*>(p), u = *reinterpret_cast<const float4*>(p + 4);
        v[0] = t.x, v[1] = t.y, v[2] = t.z, v[3] = t.w, v[4] = u.x, v[5] = u.y, v[6] = u.z, v[7] = u.w;
    } else if constexpr (R == 4) {
        const float4 t = *reinterpret_cast<const float4*>(p);
        v[0] = t.x, v[1] = t.y, v[2] = t.z, v[3] = t.w;
    } else if constexpr (R == 2) {
        const float2 t = *reinterpret_cast<const float2*>(p);
        v[0] = t.x, v[1] = t.y;
    } else {
        v[0] = p[0];
    }
}
template <int R>
__device__ __forceinline__ void store_vec(float* p, const float (&v)[R]) {
    if constexpr (R == 8) {
        *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
        *reinterpret_cast<float4*>(p + 4) = make_float4(v[4], v[5], v[6], v[7]);
    } else if constexpr (R == 4) {
        *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
    } else if constexpr (R == 2) {
        *reinterpret_cast<float2*>(p) = make_float2(v[0], v[1]);
    } else {
        p[0] = v[0];
    }
}

// x[c .. c+R-1] (all inside x): 16-byte loads at 4-byte alignment, one per
// four values (a wave then reads each diagonal's x run with R / 4
// instructions instead of R), plain loads for R < 4.
template <int R>
__device__ __forceinline__ void load_x_run(const float* __restrict__ x, int c, float (&v)[R]) {
    if constexpr (R % 4 == 0) {
#pragma unroll
        for (int q = 0; q < R / 4; ++q) {
            const f32x4a4 t = *reinterpret_cast<const f32x4a4*>(x + c + 4 * q);
            v[4 * q] = t.x, v[4 * q + 1] = t.y, v[4 * q + 2] = t.z, v[4 * q + 3] = t.w;
        }
    } else {
#pragma unroll
        for (int j = 0; j < R; ++j) v[j] = x[c + j];
    }
}

// Rows per lane of the DIA kernel and 32-bit words of values it loads per
// diagonal: 1 word for BF16 hi+lo and for 4 FP8 values, 2 words (one
// dwordx2) for 4 FP16 or int16 values.  More rows per lane means fewer
// waves: at 8 rows per lane a 128^3 stencil had 4 waves per SIMD and ran
// slower than at 2.
template <int CODEC>
constexpr int dia_rows_per_lane() {
    return CODEC == kCodecFp8Slice ? 4 : codec_row_scaled(CODEC) ? 4 : 1;
}
template <int CODEC>
constexpr int dia_words_per_lane() {
    return codec_row_scaled(CODEC) ? 2 : 1;
}

// The R values of rows row .. row+R-1 in W words, decoded; FP16 and int16
// values are returned unscaled (the row scale is applied to the row sums),
// FP8 values times the slice scale s.
template <int CODEC, int R, int W>
__device__ __forceinline__ void dia_decode(const unsigned (&w)[W], float s, float (&v)[R]) {
    if constexpr (CODEC == kCodecBf16x2) {
        v[0] = bf16x2_value(w[0]);
    } else if constexpr (CODEC == kCodecFp16Row) {
#pragma unroll
        for (int i = 0; i < W; ++i) {
            v[2 * i] = f16_value((uint16_t)(w[i] & 0xffffu));
            v[2 * i + 1] = f16_value((uint16_t)(w[i] >> 16));
        }
    } else if constexpr (CODEC == kCodecI16Row) {
#pragma unroll
        for (int i = 0; i < W; ++i) {
            v[2 * i] = i16_lo(w[i]);
            v[2 * i + 1] = i16_hi(w[i]);
        }
    } else {
#pragma unroll
        for (int j = 0; j < R; ++j) v[j] = s * e4m3_value((w[0] >> (8 * j)) & 0xffu);
    }
}

// R = dia_rows_per_lane consecutive rows per lane (a wave's value loads per
// diagonal: 256 contiguous bytes, 1 KiB for FP16 / int16) and R consecutive
// x values per diagonal.  A workgroup covers 256 R
// rows; when every column they can touch (its rows plus the smallest and
// largest offset) lies inside x -- all but the first and last workgroups of
// a banded matrix -- it takes the branch-free interior loop, chosen once per
// workgroup.  With R = 4 a wave's 256 rows are one FP8 slice (a scalar
// scale load per diagonal).
template <int CODEC>
__global__ __launch_bounds__(256) void dia_mixed_kernel(const int* __restrict__ offsets, const void* __restrict__ data,
                                                        const float* __restrict__ scale, int L, int ndiag, int nsl,
                                                        int offmin, int offmax, const float* __restrict__ x,
                                                        float* __restrict__ y, int rows, int cols) {
    constexpr int R = dia_rows_per_lane<CODEC>();
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int row = (tile * 256 + threadIdx.x) * R;
    const long long wrow0 = (long long)tile * 256 * R;
    const bool interior = wrow0 + offmin >= 0 && wrow0 + 256 * R - 1 + offmax < cols;
    if (row >= rows) return;
    const unsigned* words = reinterpret_cast<const unsigned*>(data);
    const int sl = __builtin_amdgcn_readfirstlane(row >> 8);
    auto slice_scale = [&](int d) { return CODEC == kCodecFp8Slice ? scale[(size_t)d * nsl + sl] : 1.f; };
    constexpr int W = dia_words_per_lane<CODEC>();
    // words of diagonal d for this lane's rows: a word holds R / W values, so
    // value d L + row is word (d L + row) / R * W (L and row are multiples of
    // R; the W words are 4 W bytes aligned)
    auto word = [&](int d, unsigned (&w)[W]) {
        const size_t i = ((size_t)d * L + row) / R * W;
        if constexpr (W == 4) {
            const uint4 t = *reinterpret_cast<const uint4*>(words + i);
            w[0] = t.x, w[1] = t.y, w[2] = t.z, w[3] = t.w;
        } else if constexpr (W == 2) {
            const uint2 t = *reinterpret_cast<const uint2*>(words + i);
            w[0] = t.x, w[1] = t.y;
        } else if constexpr (CODEC == kCodecBf16x2) {
            w[0] = ld_stream(words + i);
        } else {
            w[0] = words[i];
        }
    };
    // y is read before the diagonals (its latency overlaps theirs), R values
    // at once unless the lane holds the end of y
    const bool full = row + R <= rows;
    float yv[R];
    if (full) {
        load_vec<R>(y + row, yv);
    } else {
#pragma unroll
        for (int j = 0; j < R; ++j) yv[j] = row + j < rows ? y[row + j] : 0.f;
    }
    float a[R];
#pragma unroll
    for (int j = 0; j < R; ++j) a[j] = 0.f;
    // U diagonals' loads in flight per step, the last step's missing ones
    // skipped by a wave-uniform test (no serial remainder loop); the sums
    // keep diagonal order.  Boundary workgroups take the same steps with
    // bounds-tested x loads: a one-diagonal-at-a-time loop there made those
    // few workgroups the kernel's tail.
    constexpr int U = R >= 4 ? 6 : 8;
    for (int d = 0; d < ndiag; d += U) {
        unsigned w[U][W];
        float xv[U][R];
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (d + u < ndiag) {
                word(d + u, w[u]);
                const int c = row + offsets[d + u];
                if (interior) {
                    load_x_run<R>(x, c, xv[u]);
                } else {
#pragma unroll
                    for (int j = 0; j < R; ++j) xv[u][j] = xload_m(x, c + j, cols);
                }
            }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (d + u < ndiag) {
                float v[R];
                dia_decode<CODEC, R, W>(w[u], slice_scale(d + u), v);
#pragma unroll
                for (int j = 0; j < R; ++j) a[j] = fmaf(v[j], xv[u][j], a[j]);
            }
    }
    float rs[R];
    if (codec_row_scaled(CODEC)) load_vec<R>(scale + row, rs);  // the scale array has L >= row + R entries
#pragma unroll
    for (int j = 0; j < R; ++j) yv[j] += codec_row_scaled(CODEC) ? a[j] * rs[j] : a[j];
    if (full) {
        store_vec<R>(y + row, yv);
    } else {
#pragma unroll
        for (int j = 0; j < R; ++j)
            if (row + j < rows) y[row + j] = yv[j];
    }
}

// Position of value e of the block in slot idx (slots come in steps of 64,
// one per lane of a slice's wave): the BS values of a step are stored one
// after another, 64 of each, so that a wave's value loads of one step read
// one contiguous run of 64 x BS elements rather than BS streams far apart
// (HBM reads such streams ~20 % slower).
template <int BS>
__device__ __forceinline__ size_t bsr_vpos(int idx, int e) {
    return ((size_t)(idx >> 6) * BS + e) * 64 + (idx & 63);
}

// Value e of the block in slot idx, decoded (FP8: times the scale of element
// e in slice s; FP16, int16: the row scale is applied by the caller).
template <int CODEC>
__device__ __forceinline__ float bsr_value(const void* __restrict__ data, const float* __restrict__ scale, size_t i,
                                          int e, int nsl, int s) {
    if (CODEC == 0) return ld_stream(reinterpret_cast<const float*>(data) + i);  // plain FP32
    if (CODEC == kCodecBf16x2) return bf16x2_value(ld_stream(reinterpret_cast<const uint32_t*>(data) + i));
    if (CODEC == kCodecFp16Row) return f16_value(ld_narrow(reinterpret_cast<const uint16_t*>(data) + i));
    if (CODEC == kCodecI16Row) return (float)ld_narrow(reinterpret_cast<const int16_t*>(data) + i);
    return scale[(size_t)e * nsl + s] * e4m3_value(ld_narrow(reinterpret_cast<const uint8_t*>(data) + i));
}

// As bsr_sliced_kernel (blocks G at a time, loads issued back to back, x
// through load_xblock or, XW, from the workgroup's x window staged in LDS
// with the first G blocks' loads issued ahead of the staging; cols >= BW),
// with encoded values and optionally int16 column steps: the G steps are
// loaded together and summed in order.
template <int BH, int BW, int CODEC, bool DELTA, bool XW>
__global__ __launch_bounds__(XW ? kBsrWinRows : 256) void bsr_mixed_kernel(
    const int* __restrict__ sp, const void* __restrict__ colv, const int* __restrict__ base,
    const void* __restrict__ data, const float* __restrict__ scale, int stride, int brows, int nsl,
    const int2* __restrict__ win, const float* __restrict__ x, float* __restrict__ y, int rows, int cols) {
    constexpr int G = bsr_group(BH * BW);
    constexpr int T = XW ? kBsrWinRows : 256;
    const int wg = xcd_tile(blockIdx.x, gridDim.x);
    const int r = wg * T + threadIdx.x;
    __shared__ __align__(16) float xs[XW ? kBsrWinFloats : 4];
    int w0 = 0;
    auto xblock = [&](int c, float(&xv)[BW]) {
        if constexpr (XW) {
#pragma unroll
            for (int j = 0; j < BW; ++j) xv[j] = xs[c - w0 + j];
        } else {
            load_xblock<BW>(x, c, cols, xv);
        }
    };
    const bool live = r < brows;
    if (!XW && !live) return;
    // a wave's 64 block rows are one slice: wave-uniform slice quantities
    // (waves past the last block row have none and only stage x)
    const int s = __builtin_amdgcn_readfirstlane(r >> 6), t = r & 63;
    const int b = live ? sp[s] : 0;
    const int width = live ? (sp[s + 1] - b) >> 6 : 0;
    float acc[BH];
#pragma unroll
    for (int h = 0; h < BH; ++h) acc[h] = 0.f;
    int cprev = DELTA && live ? base[r] : 0;
    auto load_col = [&](int idx) {
        if (DELTA) return (int)ld_narrow(reinterpret_cast<const int16_t*>(colv) + idx);
        return ld_stream(reinterpret_cast<const int*>(colv) + idx);
    };
    // columns and values of blocks k .. k+G-1
    auto load_group = [&](int k, int (&c)[G], float (&v)[G][BH * BW]) {
#pragma unroll
        for (int u = 0; u < G; ++u) c[u] = load_col(b + ((k + u) << 6) + t);
#pragma unroll
        for (int u = 0; u < G; ++u)
#pragma unroll
            for (int e = 0; e < BH * BW; ++e)
                v[u][e] = bsr_value<CODEC>(data, scale, bsr_vpos<BH * BW>(b + ((k + u) << 6) + t, e), e, nsl, s);
    };
    auto do_group = [&](int (&c)[G], float (&v)[G][BH * BW]) {
        if (DELTA) {
#pragma unroll
            for (int u = 0; u < G; ++u) {
                cprev += c[u];
                c[u] = cprev;
            }
        }
        float xv[G][BW];
#pragma unroll
        for (int u = 0; u < G; ++u) xblock(c[u], xv[u]);
#pragma unroll
        for (int u = 0; u < G; ++u)
#pragma unroll
            for (int h = 0; h < BH; ++h)
#pragma unroll
                for (int w = 0; w < BW; ++w) acc[h] = fmaf(v[u][h * BW + w], xv[u][w], acc[h]);
    };
    int k = 0;
    if constexpr (XW) {
        int c[G];
        float v[G][BH * BW];
        const bool first = width >= G;
        if (first) load_group(0, c, v);
        const int2 w = win[wg];
        w0 = w.x;
        stage_xwindow<T>(x, w.x, w.y, cols, xs);
        __syncthreads();
        if (!live) return;
        if (first) {
            do_group(c, v);
            k = G;
        }
    }
    for (; k + G <= width; k += G) {
        int c[G];
        float v[G][BH * BW];
        load_group(k, c, v);
        do_group(c, v);
    }
    for (; k < width; ++k) {
        const int idx = b + (k << 6) + t;
        int c = load_col(idx);
        if (DELTA) {
            cprev += c;
            c = cprev;
        }
        float xv[BW];
        xblock(c, xv);
#pragma unroll
        for (int h = 0; h < BH; ++h)
#pragma unroll
            for (int w = 0; w < BW; ++w)
                acc[h] = fmaf(bsr_value<CODEC>(data, scale, bsr_vpos<BH * BW>(idx, h * BW + w), h * BW + w, nsl, s),
                              xv[w], acc[h]);
    }
    const float rs = codec_row_scaled(CODEC) ? scale[r] : 1.f;
#pragma unroll
    for (int h = 0; h < BH; ++h) {
        const int row = r * BH + h;
        if (row < rows) y[row] += codec_row_scaled(CODEC) ? rs * acc[h] : acc[h];
    }
}

}  // namespace dev

int launch_dia_mixed(const DevDiaMixed& m, const float* x, float* y, hipStream_t st) {
    if (m.rows == 0) return 0;
#define CLSPMV_DIA_MIXED(C)                                                                                     \
    hipLaunchKernelGGL(dev::dia_mixed_kernel<C>, dim3((m.rows + 256 * dev::dia_rows_per_lane<C>() - 1) /      \
                                                      (256 * dev::dia_rows_per_lane<C>())),                   \
                       dim3(256), 0, st, m.offsets, m.data, m.scale, m.L, m.ndiag, m.nsl, m.offmin, m.offmax, x, \
                       y, m.rows, m.cols)
    switch (m.codec) {
        case kCodecBf16x2: CLSPMV_DIA_MIXED(kCodecBf16x2); break;
        case kCodecFp16Row: CLSPMV_DIA_MIXED(kCodecFp16Row); break;
        case kCodecI16Row: CLSPMV_DIA_MIXED(kCodecI16Row); break;
        default: CLSPMV_DIA_MIXED(kCodecFp8Slice); break;
    }
#undef CLSPMV_DIA_MIXED
    return (int)hipGetLastError();
}

template <int BH, int BW, bool XW>
static void bsr_mixed_launch(const DevBsrMixed& m, const float* x, float* y, hipStream_t st) {
    constexpr int T = XW ? kBsrWinRows : 256;
    const dim3 g((m.brows + T - 1) / T), b(T);
#define CLSPMV_BSR_MIXED(C, D)                                                                                     \
    hipLaunchKernelGGL((dev::bsr_mixed_kernel<BH, BW, C, D, XW>), g, b, 0, st, m.sp, m.col, m.base, m.data,       \
                       m.scale, m.stride, m.brows, m.nsl, m.win, x, y, m.rows, m.cols)
    if (m.delta) {
        if (m.codec == 0) CLSPMV_BSR_MIXED(0, true);
        else if (m.codec == kCodecBf16x2) CLSPMV_BSR_MIXED(kCodecBf16x2, true);
        else if (m.codec == kCodecFp16Row) CLSPMV_BSR_MIXED(kCodecFp16Row, true);
        else if (m.codec == kCodecI16Row) CLSPMV_BSR_MIXED(kCodecI16Row, true);
        else CLSPMV_BSR_MIXED(kCodecFp8Slice, true);
    } else {
        if (m.codec == 0) CLSPMV_BSR_MIXED(0, false);
        else if (m.codec == kCodecBf16x2) CLSPMV_BSR_MIXED(kCodecBf16x2, false);
        else if (m.codec == kCodecFp16Row) CLSPMV_BSR_MIXED(kCodecFp16Row, false);
        else if (m.codec == kCodecI16Row) CLSPMV_BSR_MIXED(kCodecI16Row, false);
        else CLSPMV_BSR_MIXED(kCodecFp8Slice, false);
    }
#undef CLSPMV_BSR_MIXED
}

int launch_bsr_mixed(const DevBsrMixed& m, const float* x, float* y, hipStream_t st) {
    if (m.brows == 0) return 0;
    switch (m.bh * 10 + m.bw + (m.win ? 100 : 0)) {
        case 22: bsr_mixed_launch<2, 2, false>(m, x, y, st); break;
        case 33: bsr_mixed_launch<3, 3, false>(m, x, y, st); break;
        case 44: bsr_mixed_launch<4, 4, false>(m, x, y, st); break;
        case 122: bsr_mixed_launch<2, 2, true>(m, x, y, st); break;
        case 133: bsr_mixed_launch<3, 3, true>(m, x, y, st); break;
        case 144: bsr_mixed_launch<4, 4, true>(m, x, y, st); break;
        default: return (int)hipErrorInvalidValue;
    }
    return (int)hipGetLastError();
}

}  // namespace clspmv
