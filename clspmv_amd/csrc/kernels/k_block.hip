// Blocked-format SpMV kernels (y += A*x) for gfx950: BELL (b4ell), SBELL,
// BCSR (b4csr) with the reference's float4-column layouts and block shapes
// bh in {1,2,4,8} x bw in {4,8}, plus a native r x c BSR (e.g. 3x3).
//
//  * BELL / SBELL: one lane per block row (A/kernels/spmv_bell.cl,
//    spmv_sbell.cl use the same mapping); the value layouts make lane r's
//    float4 for (slot, w4, h) adjacent to lane r+1's, so each dwordx4 load
//    of a wave reads 1 KiB of contiguous HBM.  x is read as float4 at the
//    block's float4 column (guarded at the right matrix edge).
//  * BCSR: the reference gives every block row a 128-thread workgroup with an
//    LDS tree reduction (A/kernels/spmv_bcsr.cl:4-44).  Here a group of L
//    lanes (L from the mean blocks per block row) strides the blocks of one
//    block row and reduces with shuffles, so short block rows do not idle a
//    whole workgroup.
//  * BSR r x c: element-major values data[e*stride + i]; same L-lane scheme.
#include "kcommon.hpp"
#include "launch.hpp"

namespace clspmv {
namespace dev {

// x as float4 at float4-column c4 (elements 4*c4 .. 4*c4+3), zero past `cols`.
__device__ __forceinline__ float4 x4load(const float* __restrict__ x, int c4, int cols) {
    const int c = c4 * 4;
    if (c + 3 < cols) return *reinterpret_cast<const float4*>(x + c);
    float4 r;
    r.x = (c < cols) ? x[c] : 0.f;
    r.y = (c + 1 < cols) ? x[c + 1] : 0.f;
    r.z = (c + 2 < cols) ? x[c + 2] : 0.f;
    r.w = 0.f;
    return r;
}

__device__ __forceinline__ float dot4(float4 a, float4 b, float acc) {
    acc = fmaf(a.x, b.x, acc);
    acc = fmaf(a.y, b.y, acc);
    acc = fmaf(a.z, b.z, acc);
    return fmaf(a.w, b.w, acc);
}

template <int BH, int W4>
__global__ __launch_bounds__(256) void bell_kernel(const int* __restrict__ col, const float4* __restrict__ data,
                                                   int Hc, int F4q, int slots, int brows,
                                                   const float* __restrict__ x, float* __restrict__ y, int rows,
                                                   int cols) {
    const int r = xcd_tile(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    if (r >= brows) return;
    float acc[BH];
#pragma unroll
    for (int h = 0; h < BH; ++h) acc[h] = 0.f;
    for (int j = 0; j < slots; ++j) {
        const int c4 = ld_stream(col + (size_t)j * Hc + r);
        float4 xv[W4];
#pragma unroll
        for (int w4 = 0; w4 < W4; ++w4) xv[w4] = x4load(x, c4 + w4, cols);
        const float4* dp = data + (size_t)j * (W4 * BH) * F4q + r;
#pragma unroll
        for (int w4 = 0; w4 < W4; ++w4)
#pragma unroll
            for (int h = 0; h < BH; ++h) acc[h] = dot4(ld_stream(dp + (size_t)(w4 * BH + h) * F4q), xv[w4], acc[h]);
    }
#pragma unroll
    for (int h = 0; h < BH; ++h) {
        const int row = r * BH + h;
        if (row < rows) y[row] += acc[h];
    }
}

template <int BH, int W4>
__global__ __launch_bounds__(256) void sbell_kernel(const int* __restrict__ slice_ptr, const int* __restrict__ col,
                                                    const float4* __restrict__ data, int S, int brows,
                                                    const float* __restrict__ x, float* __restrict__ y, int rows,
                                                    int cols) {
    const int r = xcd_tile(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    if (r >= brows) return;
    const int s = r / S, t = r - s * S;
    const int b = slice_ptr[s];
    const int width = (slice_ptr[s + 1] - b) / S;
    float acc[BH];
#pragma unroll
    for (int h = 0; h < BH; ++h) acc[h] = 0.f;
    // float4 index of (slot j, w4, h, t) = b*W4*BH + j*(S*W4*BH) + w4*(S*BH) + h*S + t
    const float4* dbase = data + (size_t)b * (W4 * BH) + t;
    for (int j = 0; j < width; ++j) {
        const int c4 = ld_stream(col + b + t + (size_t)j * S);
        float4 xv[W4];
#pragma unroll
        for (int w4 = 0; w4 < W4; ++w4) xv[w4] = x4load(x, c4 + w4, cols);
        const float4* dp = dbase + (size_t)j * (S * W4 * BH);
#pragma unroll
        for (int w4 = 0; w4 < W4; ++w4)
#pragma unroll
            for (int h = 0; h < BH; ++h) acc[h] = dot4(ld_stream(dp + (size_t)(w4 * BH + h) * S), xv[w4], acc[h]);
    }
#pragma unroll
    for (int h = 0; h < BH; ++h) {
        const int row = r * BH + h;
        if (row < rows) y[row] += acc[h];
    }
}

template <int BH, int W4, int L>
__global__ __launch_bounds__(256) void bcsr_kernel(const int* __restrict__ rp, const int* __restrict__ col,
                                                   const float4* __restrict__ data, int N4, int brows,
                                                   const float* __restrict__ x, float* __restrict__ y, int rows,
                                                   int cols) {
    const int gid = xcd_tile(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    const int lane = threadIdx.x & (L - 1);
    const int r = gid / L;
    if (r >= brows) return;
    float acc[BH];
#pragma unroll
    for (int h = 0; h < BH; ++h) acc[h] = 0.f;
    const int s = rp[r], e = rp[r + 1];
    for (int i = s + lane; i < e; i += L) {
        const int c4 = ld_stream(col + i);
#pragma unroll
        for (int w4 = 0; w4 < W4; ++w4) {
            const float4 xv = x4load(x, c4 + w4, cols);
#pragma unroll
            for (int h = 0; h < BH; ++h) acc[h] = dot4(ld_stream(data + (size_t)(w4 * BH + h) * N4 + i), xv, acc[h]);
        }
    }
#pragma unroll
    for (int h = 0; h < BH; ++h) acc[h] = group_sum<L>(acc[h]);
    if (lane == 0 && e > s) {
#pragma unroll
        for (int h = 0; h < BH; ++h) {
            const int row = r * BH + h;
            if (row < rows) y[row] += acc[h];
        }
    }
}

template <int BH, int BW, int L>
__global__ __launch_bounds__(256) void bsr_kernel(const int* __restrict__ rp, const int* __restrict__ col,
                                                  const float* __restrict__ data, int stride, int brows,
                                                  const float* __restrict__ x, float* __restrict__ y, int rows,
                                                  int cols) {
    const int gid = xcd_tile(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    const int lane = threadIdx.x & (L - 1);
    const int r = gid / L;
    if (r >= brows) return;
    float acc[BH];
#pragma unroll
    for (int h = 0; h < BH; ++h) acc[h] = 0.f;
    const int s = rp[r], e = rp[r + 1];
    for (int i = s + lane; i < e; i += L) {
        const int c = ld_stream(col + i);
        float xv[BW];
#pragma unroll
        for (int w = 0; w < BW; ++w) xv[w] = (c + w < cols) ? x[c + w] : 0.f;
#pragma unroll
        for (int h = 0; h < BH; ++h)
#pragma unroll
            for (int w = 0; w < BW; ++w) acc[h] = fmaf(ld_stream(data + (size_t)(h * BW + w) * stride + i), xv[w], acc[h]);
    }
#pragma unroll
    for (int h = 0; h < BH; ++h) acc[h] = group_sum<L>(acc[h]);
    if (lane == 0 && e > s) {
#pragma unroll
        for (int h = 0; h < BH; ++h) {
            const int row = r * BH + h;
            if (row < rows) y[row] += acc[h];
        }
    }
}

// Sliced BSR: slices of 64 block rows (one wave), per-slice width; the slot
// of (slice s, block row t, k) is sp[s] + k*64 + t, so lane t's loads of the
// column and of each of the BH*BW value streams are 256 contiguous bytes per
// wave.  Padding slots repeat the row's last column with zero values.
// Blocks go G = bsr_group(BH * BW) at a time: their column loads, then all
// x and value loads of the G blocks, are issued back to back (x through
// load_xblock: no load sits behind a branch).  Requires cols >= BW.
// XW: kBsrWinRows-thread workgroups first stage their x window win[] into
// LDS (the columns of a wide band: one coalesced pass instead of a gather
// of one cache line per lane and block), with the column and value loads
// of each lane's first G blocks issued ahead of the staging; each lane's
// sums are the same operations in the same order either way.
template <int BH, int BW, bool XW>
__global__ __launch_bounds__(XW ? kBsrWinRows : 256) void bsr_sliced_kernel(
    const int* __restrict__ sp, const int* __restrict__ col, const float* __restrict__ data, int stride, int brows,
    const int2* __restrict__ win, const float* __restrict__ x, float* __restrict__ y, int rows, int cols) {
    constexpr int G = bsr_group(BH * BW);
    constexpr int T = XW ? kBsrWinRows : 256;
    const int wg = xcd_tile(blockIdx.x, gridDim.x);
    const int r = wg * T + threadIdx.x;
    __shared__ __align__(16) float xs[XW ? kBsrWinFloats : 4];
    auto xblock = [&](int w0, int c, float(&xv)[BW]) {
        if constexpr (XW) {
#pragma unroll
            for (int j = 0; j < BW; ++j) xv[j] = xs[c - w0 + j];
        } else {
            load_xblock<BW>(x, c, cols, xv);
        }
    };
    auto load_group = [&](int b, int k, int t, int (&c)[G], float (&v)[G][BH * BW]) {
#pragma unroll
        for (int u = 0; u < G; ++u) c[u] = ld_stream(col + b + ((k + u) << 6) + t);
#pragma unroll
        for (int u = 0; u < G; ++u)
#pragma unroll
            for (int e = 0; e < BH * BW; ++e) v[u][e] = ld_stream(data + (size_t)e * stride + b + ((k + u) << 6) + t);
    };
    const bool live = r < brows;
    if (!XW && !live) return;
    const int s = r >> 6, t = r & 63;
    const int b = live ? sp[s] : 0;
    const int width = live ? (sp[s + 1] - b) >> 6 : 0;
    float acc[BH];
#pragma unroll
    for (int h = 0; h < BH; ++h) acc[h] = 0.f;
    int w0 = 0;
    int k = 0;
    if constexpr (XW) {
        int c[G];
        float v[G][BH * BW];
        const bool first = width >= G;
        if (first) load_group(b, 0, t, c, v);
        const int2 w = win[wg];
        w0 = w.x;
        stage_xwindow<T>(x, w.x, w.y, cols, xs);
        __syncthreads();
        if (!live) return;
        if (first) {
            float xv[G][BW];
#pragma unroll
            for (int u = 0; u < G; ++u) xblock(w0, c[u], xv[u]);
#pragma unroll
            for (int u = 0; u < G; ++u)
#pragma unroll
                for (int h = 0; h < BH; ++h)
#pragma unroll
                    for (int q = 0; q < BW; ++q) acc[h] = fmaf(v[u][h * BW + q], xv[u][q], acc[h]);
            k = G;
        }
    }
    for (; k + G <= width; k += G) {
        int c[G];
#pragma unroll
        for (int u = 0; u < G; ++u) c[u] = ld_stream(col + b + ((k + u) << 6) + t);
        float xv[G][BW], v[G][BH * BW];
#pragma unroll
        for (int u = 0; u < G; ++u) xblock(w0, c[u], xv[u]);
#pragma unroll
        for (int u = 0; u < G; ++u)
#pragma unroll
            for (int e = 0; e < BH * BW; ++e) v[u][e] = ld_stream(data + (size_t)e * stride + b + ((k + u) << 6) + t);
#pragma unroll
        for (int u = 0; u < G; ++u)
#pragma unroll
            for (int h = 0; h < BH; ++h)
#pragma unroll
                for (int w = 0; w < BW; ++w) acc[h] = fmaf(v[u][h * BW + w], xv[u][w], acc[h]);
    }
    for (; k < width; ++k) {
        const int idx = b + (k << 6) + t;
        float xv[BW];
        xblock(w0, ld_stream(col + idx), xv);
#pragma unroll
        for (int h = 0; h < BH; ++h)
#pragma unroll
            for (int w = 0; w < BW; ++w) acc[h] = fmaf(ld_stream(data + (size_t)(h * BW + w) * stride + idx), xv[w], acc[h]);
    }
#pragma unroll
    for (int h = 0; h < BH; ++h) {
        const int row = r * BH + h;
        if (row < rows) y[row] += acc[h];
    }
}

// Plain copy of n4 float4 (the measured streaming rate): four independent
// 16-byte loads per thread in flight, one workgroup per 1024 float4.
__global__ __launch_bounds__(256) void copy_probe_kernel(const float4* __restrict__ src, float4* __restrict__ dst,
                                                         size_t n4) {
    const size_t base = (size_t)blockIdx.x * 1024 + threadIdx.x;
    float4 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const size_t i = base + (size_t)u * 256;
        v[u] = i < n4 ? ld_stream(src + i) : make_float4(0.f, 0.f, 0.f, 0.f);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const size_t i = base + (size_t)u * 256;
        if (i < n4) dst[i] = v[u];
    }
}

__global__ void empty_kernel() {}

// Counts the 16-byte words where a and b differ (the copy probe's check).
__global__ __launch_bounds__(256) void compare_probe_kernel(const uint4* __restrict__ a, const uint4* __restrict__ b,
                                                            size_t n4, unsigned long long* __restrict__ count) {
    unsigned long long bad = 0;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n4; i += (size_t)gridDim.x * 256) {
        const uint4 u = a[i], v = b[i];
        bad += (u.x != v.x) | (u.y != v.y) | (u.z != v.z) | (u.w != v.w);
    }
    if (bad) atomicAdd(count, bad);
}

}  // namespace dev

#define BH_W4_SWITCH(BH, W4, MACRO)     \
    switch (BH * 10 + W4) {             \
        case 11: MACRO(1, 1); break;    \
        case 21: MACRO(2, 1); break;    \
        case 41: MACRO(4, 1); break;    \
        case 81: MACRO(8, 1); break;    \
        case 12: MACRO(1, 2); break;    \
        case 22: MACRO(2, 2); break;    \
        case 42: MACRO(4, 2); break;    \
        case 82: MACRO(8, 2); break;    \
        default: return (int)hipErrorInvalidValue; \
    }

int launch_bell(const DevBell& m, const float* x, float* y, hipStream_t st) {
    if (m.brows == 0) return 0;
    dim3 grid((m.brows + 255) / 256);
#define L_BELL(BH, W4)                                                                                        \
    hipLaunchKernelGGL((dev::bell_kernel<BH, W4>), grid, dim3(256), 0, st, m.col,                             \
                       reinterpret_cast<const float4*>(m.data), m.Hc, m.F4 / 4, m.slots, m.brows, x, y, m.rows, \
                       m.cols)
    BH_W4_SWITCH(m.bh, m.w4n, L_BELL)
#undef L_BELL
    return (int)hipGetLastError();
}

int launch_sbell(const DevSbell& m, const float* x, float* y, hipStream_t st) {
    if (m.brows == 0) return 0;
    dim3 grid((m.brows + 255) / 256);
#define L_SBELL(BH, W4)                                                                                          \
    hipLaunchKernelGGL((dev::sbell_kernel<BH, W4>), grid, dim3(256), 0, st, m.slice_ptr, m.col,                  \
                       reinterpret_cast<const float4*>(m.data), m.S, m.brows, x, y, m.rows, m.cols)
    BH_W4_SWITCH(m.bh, m.w4n, L_SBELL)
#undef L_SBELL
    return (int)hipGetLastError();
}

template <int BH, int W4>
static void launch_bcsr_l(const DevBcsr& m, const float* x, float* y, hipStream_t st) {
    const long long threads = (long long)m.brows * m.lanes;
    dim3 grid((unsigned)((threads + 255) / 256));
    const float4* d4 = reinterpret_cast<const float4*>(m.data);
    switch (m.lanes) {
        case 2: hipLaunchKernelGGL((dev::bcsr_kernel<BH, W4, 2>), grid, dim3(256), 0, st, m.row_ptr, m.col, d4, m.N4, m.brows, x, y, m.rows, m.cols); break;
        case 4: hipLaunchKernelGGL((dev::bcsr_kernel<BH, W4, 4>), grid, dim3(256), 0, st, m.row_ptr, m.col, d4, m.N4, m.brows, x, y, m.rows, m.cols); break;
        case 8: hipLaunchKernelGGL((dev::bcsr_kernel<BH, W4, 8>), grid, dim3(256), 0, st, m.row_ptr, m.col, d4, m.N4, m.brows, x, y, m.rows, m.cols); break;
        default: hipLaunchKernelGGL((dev::bcsr_kernel<BH, W4, 16>), grid, dim3(256), 0, st, m.row_ptr, m.col, d4, m.N4, m.brows, x, y, m.rows, m.cols); break;
    }
}

int launch_bcsr(const DevBcsr& m, const float* x, float* y, hipStream_t st) {
    if (m.brows == 0) return 0;
#define L_BCSR(BH, W4) launch_bcsr_l<BH, W4>(m, x, y, st)
    BH_W4_SWITCH(m.bh, m.w4n, L_BCSR)
#undef L_BCSR
    return (int)hipGetLastError();
}

template <int BH, int BW>
static void launch_bsr_l(const DevBsr& m, const float* x, float* y, hipStream_t st) {
    const long long threads = (long long)m.brows * m.lanes;
    dim3 grid((unsigned)((threads + 255) / 256));
    switch (m.lanes) {
        case 2: hipLaunchKernelGGL((dev::bsr_kernel<BH, BW, 2>), grid, dim3(256), 0, st, m.row_ptr, m.col, m.data, m.stride, m.brows, x, y, m.rows, m.cols); break;
        case 4: hipLaunchKernelGGL((dev::bsr_kernel<BH, BW, 4>), grid, dim3(256), 0, st, m.row_ptr, m.col, m.data, m.stride, m.brows, x, y, m.rows, m.cols); break;
        case 8: hipLaunchKernelGGL((dev::bsr_kernel<BH, BW, 8>), grid, dim3(256), 0, st, m.row_ptr, m.col, m.data, m.stride, m.brows, x, y, m.rows, m.cols); break;
        default: hipLaunchKernelGGL((dev::bsr_kernel<BH, BW, 16>), grid, dim3(256), 0, st, m.row_ptr, m.col, m.data, m.stride, m.brows, x, y, m.rows, m.cols); break;
    }
}

template <int BH, int BW>
static void launch_bsr_sliced(const DevBsr& m, const float* x, float* y, hipStream_t st) {
    if (m.win)
        hipLaunchKernelGGL((dev::bsr_sliced_kernel<BH, BW, true>), dim3((m.brows + kBsrWinRows - 1) / kBsrWinRows),
                           dim3(kBsrWinRows), 0, st, m.row_ptr, m.col, m.data, m.stride, m.brows, m.win, x, y, m.rows,
                           m.cols);
    else
        hipLaunchKernelGGL((dev::bsr_sliced_kernel<BH, BW, false>), dim3((m.brows + 255) / 256), dim3(256), 0, st,
                           m.row_ptr, m.col, m.data, m.stride, m.brows, m.win, x, y, m.rows, m.cols);
}

int launch_bsr(const DevBsr& m, const float* x, float* y, hipStream_t st) {
    if (m.brows == 0) return 0;
    if (m.lanes == 0) {
        // 2x2..4x4 sliced parts are built with the coded layout (int16
        // column steps) and launched by launch_bsr_mixed
        switch (m.bh * 10 + m.bw) {
            case 55: launch_bsr_sliced<5, 5>(m, x, y, st); break;
            case 66: launch_bsr_sliced<6, 6>(m, x, y, st); break;
            default: return (int)hipErrorInvalidValue;
        }
        return (int)hipGetLastError();
    }
    switch (m.bh * 10 + m.bw) {
        case 22: launch_bsr_l<2, 2>(m, x, y, st); break;
        case 33: launch_bsr_l<3, 3>(m, x, y, st); break;
        case 44: launch_bsr_l<4, 4>(m, x, y, st); break;
        case 55: launch_bsr_l<5, 5>(m, x, y, st); break;
        case 66: launch_bsr_l<6, 6>(m, x, y, st); break;
        default: return (int)hipErrorInvalidValue;
    }
    return (int)hipGetLastError();
}

int launch_copy_probe(const float4* src, float4* dst, size_t n4, hipStream_t st) {
    if (n4 == 0) return 0;
    hipLaunchKernelGGL(dev::copy_probe_kernel, dim3((unsigned)((n4 + 1023) / 1024)), dim3(256), 0, st, src, dst, n4);
    return (int)hipGetLastError();
}

int launch_compare_probe(const void* a, const void* b, size_t n4, unsigned long long* count, hipStream_t st) {
    if (n4 == 0) return 0;
    const unsigned groups = (unsigned)std::min<size_t>(2048, (n4 + 255) / 256);
    hipLaunchKernelGGL(dev::compare_probe_kernel, dim3(groups), dim3(256), 0, st, (const uint4*)a, (const uint4*)b, n4,
                       count);
    return (int)hipGetLastError();
}

int launch_empty(hipStream_t st, int groups, int threads) {
    hipLaunchKernelGGL(dev::empty_kernel, dim3(groups), dim3(threads), 0, st);
    return (int)hipGetLastError();
}

}  // namespace clspmv
