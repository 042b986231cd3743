// DIA and BDIA SpMV kernels (y += A*x) for gfx950.
//
// Both use the DIA data layout data[d*L + row] (A/include/matrix_storage.h
// :30-47), so a wave's load of one diagonal is 256 contiguous bytes.  x is
// read with a bounds test instead of the reference's zero-padded copy
// (vec_extended, A/src/spmv_cocktail.cpp:253-301).
//   DIA  method 0: one row per lane (A/kernels/spmv_dia.cl:4-28 "gpu_dia");
//        method 1: four rows per lane, dwordx4 data and x loads (x at
//                  4-byte alignment), a branch-free loop for interior
//                  workgroups ("gpu_dia_v4");
//        method 4: 2-8 lanes per row splitting the diagonals (small
//                  matrices; no reference counterpart).
//   BDIA method 0: one row per lane, diagonals of a band read x directly
//                  (A/kernels/spmv_bdia.cl:46-79 "gpu_bdia_nlvec");
//        method 1: the x window of each band (256 + width - 1 values) is
//                  staged in LDS once per workgroup and shared by the band's
//                  diagonals ("gpu_bdia", with the barrier the reference
//                  omits between bands, spmv_bdia.cl:27-41).
// Offsets are uniform across the workgroup and are read with scalar loads.
#include "kcommon.hpp"
#include "launch.hpp"

namespace clspmv {
namespace dev {

__device__ __forceinline__ float xload(const float* __restrict__ x, int c, int cols) {
    return (c >= 0 && c < cols) ? x[c] : 0.f;
}

__global__ __launch_bounds__(256) void dia_row_kernel(const int* __restrict__ offsets, const float* __restrict__ data,
                                                      int L, int ndiag, const float* __restrict__ x,
                                                      float* __restrict__ y, int rows, int cols) {
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int row = tile * 256 + threadIdx.x;
    if (row >= rows) return;
    float acc = y[row];
    const float* dp = data + row;
    // 8 diagonals' loads in flight per step, the last step's missing ones
    // skipped by a wave-uniform test: 27 diagonals take 4 dependent steps
    // (6 + 3 serial ones with steps of 4), which is what bounds a small
    // matrix; the sums keep diagonal order
    for (int d = 0; d < ndiag; d += 8) {
        float v[8], xv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            v[u] = 0.f;
            xv[u] = 0.f;
            if (d + u < ndiag) {
                v[u] = ld_stream(dp + (size_t)(d + u) * L);
                xv[u] = xload(x, row + offsets[d + u], cols);
            }
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (d + u < ndiag) acc = fmaf(v[u], xv[u], acc);
    }
    y[row] = acc;
}

// G lanes per row (method 4, for small matrices): lane t of a row's group
// takes diagonals t, t + G, ..., at most 8 of them, so all of a row's loads
// issue in one round where the one-row-per-lane kernel takes ndiag / 8
// dependent rounds -- the bound of a matrix too small to fill the GPU (a
// 27-point stencil of 10^3..30^3 rows: ~4 us one row per lane, the launch
// floor ~2.6 us).  The G partial sums are added by log2 G lane exchanges
// (fixed order: deterministic).
template <int G>
__global__ __launch_bounds__(256) void dia_lanes_kernel(const int* __restrict__ offsets,
                                                        const float* __restrict__ data, int L, int ndiag,
                                                        const float* __restrict__ x, float* __restrict__ y, int rows,
                                                        int cols) {
    const int gid = blockIdx.x * 256 + threadIdx.x;
    const int row = gid / G, t = gid % G;
    float acc = 0.f;
    if (row < rows) {
        for (int d0 = t; d0 < ndiag; d0 += 8 * G) {
            float v[8], xv[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int d = d0 + u * G;
                v[u] = 0.f;
                xv[u] = 0.f;
                if (d < ndiag) {
                    v[u] = ld_stream(data + (size_t)d * L + row);
                    xv[u] = xload(x, row + offsets[d], cols);
                }
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) acc = fmaf(v[u], xv[u], acc);
        }
    }
#pragma unroll
    for (int m = 1; m < G; m <<= 1) acc += __shfl_xor(acc, m);
    if (row < rows && t == 0) y[row] += acc;
}

// Four consecutive rows per lane: one dwordx4 value load and one dwordx4
// x load (x[c .. c+3], c = row + offset, four consecutive columns) per
// diagonal.  A workgroup covers 1024 rows; when every column it can touch
// (rows row0 .. row0+1023 plus the smallest and largest offset) lies inside
// x -- all but the first and last few workgroups of a banded matrix --
// it takes the branch-free interior loop, chosen once per workgroup.
__global__ __launch_bounds__(256) void dia_row4_kernel(const int* __restrict__ offsets,
                                                       const float4* __restrict__ data, int L4, int ndiag,
                                                       int offmin, int offmax, const float* __restrict__ x,
                                                       float* __restrict__ y, int rows, int cols) {
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int r4 = tile * 256 + threadIdx.x;
    const int row = r4 * 4;
    const long long wrow0 = (long long)tile * 1024;
    const bool interior = wrow0 + offmin >= 0 && wrow0 + 1023 + offmax < cols;
    if (row >= rows) return;
    float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
    // 6 diagonals per step (the last step's missing ones skipped, uniform
    // test); boundary workgroups take the same steps with bounds-tested x
    for (int d = 0; d < ndiag; d += 6) {
        float4 v[6];
        float xv[6][4];
#pragma unroll
        for (int u = 0; u < 6; ++u)
            if (d + u < ndiag) {
                v[u] = ld_stream(data + (size_t)(d + u) * L4 + r4);
                const int c = row + offsets[d + u];
                if (interior) {
                    const f32x4a4 t = *reinterpret_cast<const f32x4a4*>(x + c);
                    xv[u][0] = t.x, xv[u][1] = t.y, xv[u][2] = t.z, xv[u][3] = t.w;
                } else {
#pragma unroll
                    for (int j = 0; j < 4; ++j) xv[u][j] = xload(x, c + j, cols);
                }
            }
#pragma unroll
        for (int u = 0; u < 6; ++u)
            if (d + u < ndiag) {
                a0 = fmaf(v[u].x, xv[u][0], a0);
                a1 = fmaf(v[u].y, xv[u][1], a1);
                a2 = fmaf(v[u].z, xv[u][2], a2);
                a3 = fmaf(v[u].w, xv[u][3], a3);
            }
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

__global__ __launch_bounds__(256) void bdia_row_kernel(const int* __restrict__ bptr, const int* __restrict__ offsets,
                                                       const float* __restrict__ data, int L, int nbands,
                                                       const float* __restrict__ x, float* __restrict__ y, int rows,
                                                       int cols) {
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int row = tile * 256 + threadIdx.x;
    if (row >= rows) return;
    float acc = y[row];
    for (int b = 0; b < nbands; ++b) {
        const int s = bptr[b], e = bptr[b + 1];
        const int c0 = row + offsets[b];
        const float* dp = data + (size_t)s * L + row;
        for (int j = 0; j < e - s; ++j) acc = fmaf(ld_stream(dp + (size_t)j * L), xload(x, c0 + j, cols), acc);
    }
    y[row] = acc;
}

// LDS-staged x band window; requires every band width <= 257.
__global__ __launch_bounds__(256) void bdia_lds_kernel(const int* __restrict__ bptr, const int* __restrict__ offsets,
                                                       const float* __restrict__ data, int L, int nbands,
                                                       const float* __restrict__ x, float* __restrict__ y, int rows,
                                                       int cols) {
    __shared__ float win[512];
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int row0 = tile * 256;
    const int t = threadIdx.x;
    const int row = row0 + t;
    const bool live = row < rows;
    float acc = live ? y[row] : 0.f;
    for (int b = 0; b < nbands; ++b) {
        const int s = bptr[b], e = bptr[b + 1];
        const int w = e - s;
        const int c0 = row0 + offsets[b];
        __syncthreads();  // previous band's readers are done with `win`
        win[t] = xload(x, c0 + t, cols);
        if (t < w - 1) win[256 + t] = xload(x, c0 + 256 + t, cols);
        __syncthreads();
        if (live) {
            const float* dp = data + (size_t)s * L + row;
            for (int j = 0; j < w; ++j) acc = fmaf(ld_stream(dp + (size_t)j * L), win[t + j], acc);
        }
    }
    if (live) y[row] = acc;
}

}  // namespace dev

int launch_dia(const DevDia& m, int method, const float* x, float* y, hipStream_t st) {
    if (m.rows == 0) return 0;
    if (method == 4) {
        const long long threads = (long long)m.rows * (m.ndiag > 32 ? 8 : m.ndiag > 16 ? 4 : 2);
        const dim3 grid((unsigned)((threads + 255) / 256));
        if (m.ndiag > 32)
            hipLaunchKernelGGL(dev::dia_lanes_kernel<8>, grid, dim3(256), 0, st, m.offsets, m.data, m.L, m.ndiag, x, y,
                               m.rows, m.cols);
        else if (m.ndiag > 16)
            hipLaunchKernelGGL(dev::dia_lanes_kernel<4>, grid, dim3(256), 0, st, m.offsets, m.data, m.L, m.ndiag, x, y,
                               m.rows, m.cols);
        else
            hipLaunchKernelGGL(dev::dia_lanes_kernel<2>, grid, dim3(256), 0, st, m.offsets, m.data, m.L, m.ndiag, x, y,
                               m.rows, m.cols);
    } else if (method == 1 && (m.L % 4) == 0) {
        int r4 = (m.rows + 3) / 4;
        hipLaunchKernelGGL(dev::dia_row4_kernel, dim3((r4 + 255) / 256), dim3(256), 0, st, m.offsets,
                           reinterpret_cast<const float4*>(m.data), m.L / 4, m.ndiag, m.offmin, m.offmax, x, y,
                           m.rows, m.cols);
    } else {
        hipLaunchKernelGGL(dev::dia_row_kernel, dim3((m.rows + 255) / 256), dim3(256), 0, st, m.offsets, m.data, m.L,
                           m.ndiag, x, y, m.rows, m.cols);
    }
    return (int)hipGetLastError();
}

int launch_bdia(const DevBdia& m, int method, const float* x, float* y, hipStream_t st) {
    if (m.rows == 0) return 0;
    if (method == 1) {
        hipLaunchKernelGGL(dev::bdia_lds_kernel, dim3((m.rows + 255) / 256), dim3(256), 0, st, m.bptr, m.offsets,
                           m.data, m.L, m.nbands, x, y, m.rows, m.cols);
    } else {
        hipLaunchKernelGGL(dev::bdia_row_kernel, dim3((m.rows + 255) / 256), dim3(256), 0, st, m.bptr, m.offsets,
                           m.data, m.L, m.nbands, x, y, m.rows, m.cols);
    }
    return (int)hipGetLastError();
}

}  // namespace clspmv
