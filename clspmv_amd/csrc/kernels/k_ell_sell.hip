// ELL and SELL SpMV kernels (y += A*x), hand-written for gfx950.
//
// ELL (A/kernels/spmv_ell.cl:4-40 computes the same product): slot-major
// arrays col[slot*H + row], val[slot*H + row].  One lane per row (method 0)
// or four consecutive rows per lane with dwordx4 loads (method 1), so every
// load instruction of a wave covers 256 B (method 0) or 1 KiB (method 1) of
// contiguous HBM.  The slot loop is unrolled so that several independent
// x gathers are in flight per lane.
//
// SELL (A/kernels/spmv_sell.cl:4-58): per-slice width; element (row, slot)
// at slice_ptr[s] + slot*S + row%S.  One lane per row for any slice height S;
// for S=32 a wave64 serves two slices.
#include "kcommon.hpp"
#include "launch.hpp"

namespace clspmv {
namespace dev {

__global__ __launch_bounds__(256) void ell_row_kernel(const int* __restrict__ col, const float* __restrict__ val, int H,
                                                      int ell_num, const float* __restrict__ x,
                                                      float* __restrict__ y, int rows) {
    const int row = xcd_tile(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    if (row >= rows) return;
    float acc = y[row];
    const int* cp = col + row;
    const float* vp = val + row;
    int i = 0;
    for (; i + 8 <= ell_num; i += 8) {
        int c[8];
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            c[u] = ld_stream(cp + (size_t)(i + u) * H);
            v[u] = ld_stream(vp + (size_t)(i + u) * H);
        }
        float xv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) xv[u] = x[c[u]];
#pragma unroll
        for (int u = 0; u < 8; ++u) acc = fmaf(v[u], xv[u], acc);
    }
    for (; i < ell_num; ++i) acc = fmaf(ld_stream(vp + (size_t)i * H), x[ld_stream(cp + (size_t)i * H)], acc);
    y[row] = acc;
}

// Four rows per lane: H is a multiple of 4 (GPU_ALIGNMENT = 32).
__global__ __launch_bounds__(256) void ell_row4_kernel(const int4* __restrict__ col, const float4* __restrict__ val,
                                                       int H4, int ell_num, const float* __restrict__ x,
                                                       float* __restrict__ y, int rows) {
    const int r4 = xcd_tile(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    const int row = r4 * 4;
    if (row >= rows) return;
    float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
    const int4* cp = col + r4;
    const float4* vp = val + r4;
    int i = 0;
    for (; i + 4 <= ell_num; i += 4) {
        int4 c[4];
        float4 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            c[u] = ld_stream(cp + (size_t)(i + u) * H4);
            v[u] = ld_stream(vp + (size_t)(i + u) * H4);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            a0 = fmaf(v[u].x, x[c[u].x], a0);
            a1 = fmaf(v[u].y, x[c[u].y], a1);
            a2 = fmaf(v[u].z, x[c[u].z], a2);
            a3 = fmaf(v[u].w, x[c[u].w], a3);
        }
    }
    for (; i < ell_num; ++i) {
        int4 c = ld_stream(cp + (size_t)i * H4);
        float4 v = ld_stream(vp + (size_t)i * H4);
        a0 = fmaf(v.x, x[c.x], a0);
        a1 = fmaf(v.y, x[c.y], a1);
        a2 = fmaf(v.z, x[c.z], a2);
        a3 = fmaf(v.w, x[c.w], a3);
    }
    if (row + 3 < rows) {
        float4 yo = *reinterpret_cast<const float4*>(y + row);
        yo.x += a0;
        yo.y += a1;
        yo.z += a2;
        yo.w += a3;
        *reinterpret_cast<float4*>(y + row) = yo;
    } else {
        y[row] += a0;
        if (row + 1 < rows) y[row + 1] += a1;
        if (row + 2 < rows) y[row + 2] += a2;
    }
}

__global__ __launch_bounds__(256) void sell_row_kernel(const int* __restrict__ slice_ptr, const int* __restrict__ col,
                                                       const float* __restrict__ val, int S,
                                                       const float* __restrict__ x, float* __restrict__ y,
                                                       int rows) {
    const int row = xcd_tile(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    if (row >= rows) return;
    const int s = row / S;
    const int b = slice_ptr[s];
    const int width = (slice_ptr[s + 1] - b) / S;
    const int* cp = col + b + (row - s * S);
    const float* vp = val + b + (row - s * S);
    float acc = y[row];
    int i = 0;
    for (; i + 4 <= width; i += 4) {
        int c[4];
        float v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            c[u] = ld_stream(cp + (size_t)(i + u) * S);
            v[u] = ld_stream(vp + (size_t)(i + u) * S);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) acc = fmaf(v[u], x[c[u]], acc);
    }
    for (; i < width; ++i) acc = fmaf(ld_stream(vp + (size_t)i * S), x[ld_stream(cp + (size_t)i * S)], acc);
    y[row] = acc;
}

}  // namespace dev

int launch_ell(const DevEll& m, int method, const float* x, float* y, hipStream_t st) {
    if (m.rows == 0) return 0;
    if (method == 1) {
        int r4 = (m.rows + 3) / 4;
        hipLaunchKernelGGL(dev::ell_row4_kernel, dim3((r4 + 255) / 256), dim3(256), 0, st,
                           reinterpret_cast<const int4*>(m.col), reinterpret_cast<const float4*>(m.val), m.H / 4,
                           m.ell_num, x, y, m.rows);
    } else {
        hipLaunchKernelGGL(dev::ell_row_kernel, dim3((m.rows + 255) / 256), dim3(256), 0, st, m.col, m.val, m.H,
                           m.ell_num, x, y, m.rows);
    }
    return (int)hipGetLastError();
}

int launch_sell(const DevSell& m, const float* x, float* y, hipStream_t st) {
    if (m.rows == 0) return 0;
    hipLaunchKernelGGL(dev::sell_row_kernel, dim3((m.rows + 255) / 256), dim3(256), 0, st, m.slice_ptr, m.col, m.val,
                       m.S, x, y, m.rows);
    return (int)hipGetLastError();
}

}  // namespace clspmv
