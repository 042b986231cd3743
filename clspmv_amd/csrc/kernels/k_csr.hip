// CSR SpMV kernels (y += A*x) for gfx950.
//
//  * csr_scalar: one lane per row (A/kernels/spmv_csr_scalar.cl:5-23 is the
//    same one-thread-per-row scheme); used by spmv_all method 1.
//  * csr_vector<L>: L lanes per row (L = 2..64 picked from the mean row
//    length), DPP/shuffle group reduction.  The reference's CSR-vector uses
//    a fixed 32-lane "warp" per row (A/kernels/spmv_csr_vector.cl:4-40).
//  * csr_adaptive: row blocks planned on the host so that every wave gets a
//    balanced amount of work whatever the row-length distribution:
//      STREAM  rows [r0,r1) with <= 512 nnz in total and no row longer than
//              64: the wave streams the block's nnz range into LDS with
//              fully coalesced loads (8 per lane), then lane j sums row j;
//      VECTOR  one row of 65..1024 nnz: the wave strides the row and
//              reduces across its 64 lanes;
//      LONG    one 1024-nnz chunk of a longer row: the wave's partial is
//              stored in FP64 and a second kernel adds the partials of each
//              long row in order (deterministic, no atomics).
#include "kcommon.hpp"
#include "launch.hpp"

namespace clspmv {
namespace dev {

__global__ __launch_bounds__(256) void csr_scalar_kernel(const int* __restrict__ rp, const int* __restrict__ col,
                                                         const float* __restrict__ val,
                                                         const float* __restrict__ x, float* __restrict__ y,
                                                         int rows) {
    const int row = blockIdx.x * 256 + threadIdx.x;
    if (row >= rows) return;
    const int s = rp[row], e = rp[row + 1];
    float acc = y[row];
    for (int k = s; k < e; ++k) acc = fmaf(val[k], x[col[k]], acc);
    y[row] = acc;
}

template <int L>
__global__ __launch_bounds__(256) void csr_vector_kernel(const int* __restrict__ rp, const int* __restrict__ col,
                                                         const float* __restrict__ val,
                                                         const float* __restrict__ x, float* __restrict__ y,
                                                         int rows) {
    const int gid = blockIdx.x * 256 + threadIdx.x;
    const int lane = threadIdx.x & (L - 1);
    const int row = gid / L;
    if (row >= rows) return;  // whole L-groups exit together (L divides 256)
    const int s = rp[row], e = rp[row + 1];
    float acc = 0.f;
    int k = s + lane;
    for (; k + 3 * L < e; k += 4 * L) {
        int c0 = ld_stream(col + k), c1 = ld_stream(col + k + L), c2 = ld_stream(col + k + 2 * L),
            c3 = ld_stream(col + k + 3 * L);
        float v0 = ld_stream(val + k), v1 = ld_stream(val + k + L), v2 = ld_stream(val + k + 2 * L),
              v3 = ld_stream(val + k + 3 * L);
        acc = fmaf(v0, x[c0], acc);
        acc = fmaf(v1, x[c1], acc);
        acc = fmaf(v2, x[c2], acc);
        acc = fmaf(v3, x[c3], acc);
    }
    for (; k < e; k += L) acc = fmaf(ld_stream(val + k), x[ld_stream(col + k)], acc);
    acc = group_sum<L>(acc);
    if (lane == 0 && e > s) y[row] += acc;
}

__global__ __launch_bounds__(256) void csr_adaptive_kernel(const int* __restrict__ rp, const int* __restrict__ col,
                                                           const float* __restrict__ val,
                                                           const int4* __restrict__ blocks, int nblocks,
                                                           const float* __restrict__ x, float* __restrict__ y,
                                                           double* __restrict__ carry) {
    __shared__ float lds[4][kCsrStreamMaxNnz];
    const int wid = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int b = blockIdx.x * 4 + wid;
    if (b >= nblocks) return;
    const int4 blk = blocks[b];
    const int r0 = blk.x, tag = blk.y, n0 = blk.z, n1 = blk.w;
    if (tag < 0 || tag - r0 == 1) {
        // VECTOR (one row) or LONG (one chunk of a long row).
        float acc = 0.f;
        int k = n0 + lane;
        for (; k + 192 < n1; k += 256) {
            int c0 = ld_stream(col + k), c1 = ld_stream(col + k + 64), c2 = ld_stream(col + k + 128),
                c3 = ld_stream(col + k + 192);
            float v0 = ld_stream(val + k), v1 = ld_stream(val + k + 64), v2 = ld_stream(val + k + 128),
                  v3 = ld_stream(val + k + 192);
            acc = fmaf(v0, x[c0], acc);
            acc = fmaf(v1, x[c1], acc);
            acc = fmaf(v2, x[c2], acc);
            acc = fmaf(v3, x[c3], acc);
        }
        for (; k < n1; k += 64) acc = fmaf(ld_stream(val + k), x[ld_stream(col + k)], acc);
        acc = wave_sum(acc);
        if (lane == 0) {
            if (tag < 0)
                carry[-tag - 1] = (double)acc;
            else if (n1 > n0)
                y[r0] += acc;
        }
        return;
    }
    // STREAM: rows [r0, tag), nnz [n0, n1), n1 - n0 <= 512.
    const int n = n1 - n0;
    float* sh = lds[wid];
#pragma unroll
    for (int u = 0; u < kCsrStreamPerLane; ++u) {
        const int k = lane + 64 * u;
        if (k < n) sh[k] = ld_stream(val + n0 + k) * x[ld_stream(col + n0 + k)];
    }
    wave_lds_sync();
    const int r = r0 + lane;
    if (r < tag) {
        const int s = rp[r] - n0, e = rp[r + 1] - n0;
        if (e > s) {
            float acc = 0.f;
            for (int k = s; k < e; ++k) acc += sh[k];
            y[r] += acc;
        }
    }
}

// One column slice of the column-sliced CSR.  Same block kinds as
// csr_adaptive_kernel; STREAM blocks take each row's entry count from one
// byte per row (`counts`) and an exclusive wave scan gives the offsets.  All
// x reads of a launch fall in one slice of columns, small enough to stay in
// L2 while the entry streams pass through with nontemporal loads.
__global__ __launch_bounds__(256) void csr_slice_kernel(const unsigned char* __restrict__ counts,
                                                        const int* __restrict__ col, const float* __restrict__ val,
                                                        const int4* __restrict__ blocks, int nblocks,
                                                        const float* __restrict__ x, float* __restrict__ y,
                                                        double* __restrict__ carry) {
    __shared__ float lds[4][kCsrStreamMaxNnz];
    const int wid = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int b = blockIdx.x * 4 + wid;
    if (b >= nblocks) return;
    const int4 blk = blocks[b];
    const int r0 = blk.x, tag = blk.y, n0 = blk.z, n1 = blk.w;
    if (tag < 0 || tag - r0 == 1) {
        float acc = 0.f;
        int k = n0 + lane;
        for (; k + 192 < n1; k += 256) {
            int c0 = ld_stream(col + k), c1 = ld_stream(col + k + 64), c2 = ld_stream(col + k + 128),
                c3 = ld_stream(col + k + 192);
            float v0 = ld_stream(val + k), v1 = ld_stream(val + k + 64), v2 = ld_stream(val + k + 128),
                  v3 = ld_stream(val + k + 192);
            acc = fmaf(v0, x[c0], acc);
            acc = fmaf(v1, x[c1], acc);
            acc = fmaf(v2, x[c2], acc);
            acc = fmaf(v3, x[c3], acc);
        }
        for (; k < n1; k += 64) acc = fmaf(ld_stream(val + k), x[ld_stream(col + k)], acc);
        acc = wave_sum(acc);
        if (lane == 0) {
            if (tag < 0)
                carry[-tag - 1] = (double)acc;
            else if (n1 > n0)
                y[r0] += acc;
        }
        return;
    }
    const int n = n1 - n0;
    float* sh = lds[wid];
#pragma unroll
    for (int u = 0; u < kCsrStreamPerLane; ++u) {
        const int k = lane + 64 * u;
        if (k < n) sh[k] = ld_stream(val + n0 + k) * x[ld_stream(col + n0 + k)];
    }
    const int r = r0 + lane;
    const int cnt = (r < tag) ? (int)counts[r] : 0;
    // exclusive scan of the counts across the wave
    int incl = cnt;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        int up = __shfl_up(incl, d, 64);
        if (lane >= d) incl += up;
    }
    const int off = incl - cnt;
    wave_lds_sync();
    if (cnt > 0) {
        float acc = 0.f;
        for (int k = off; k < off + cnt; ++k) acc += sh[k];
        y[r] += acc;
    }
}

__global__ __launch_bounds__(256) void csr_long_finalize_kernel(const int* __restrict__ long_rows,
                                                                const int* __restrict__ long_first,
                                                                const double* __restrict__ carry, int n_long,
                                                                float* __restrict__ y) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n_long) return;
    double s = 0.0;
    for (int k = long_first[i]; k < long_first[i + 1]; ++k) s += carry[k];
    const int r = long_rows[i];
    y[r] = (float)((double)y[r] + s);
}

}  // namespace dev

int launch_csr_scalar(const DevCsr& m, const float* x, float* y, hipStream_t st) {
    if (m.rows == 0) return 0;
    hipLaunchKernelGGL(dev::csr_scalar_kernel, dim3((m.rows + 255) / 256), dim3(256), 0, st, m.row_ptr, m.col, m.val,
                       x, y, m.rows);
    return (int)hipGetLastError();
}

int launch_csr_vector(const DevCsr& m, int lanes, const float* x, float* y, hipStream_t st) {
    if (m.rows == 0) return 0;
    const long long threads = (long long)m.rows * lanes;
    dim3 grid((unsigned)((threads + 255) / 256));
    switch (lanes) {
#define CASE(L)                                                                                                     \
    case L:                                                                                                         \
        hipLaunchKernelGGL(dev::csr_vector_kernel<L>, grid, dim3(256), 0, st, m.row_ptr, m.col, m.val, x, y, m.rows); \
        break;
        CASE(2) CASE(4) CASE(8) CASE(16) CASE(32) CASE(64)
#undef CASE
        default:
            return (int)hipErrorInvalidValue;
    }
    return (int)hipGetLastError();
}

int launch_csr_adaptive(const DevCsrAdaptive& m, const float* x, float* y, hipStream_t st) {
    if (m.nblocks > 0) {
        hipLaunchKernelGGL(dev::csr_adaptive_kernel, dim3((m.nblocks + 3) / 4), dim3(256), 0, st, m.csr.row_ptr,
                           m.csr.col, m.csr.val, m.blocks, m.nblocks, x, y, m.carry);
        int e = (int)hipGetLastError();
        if (e) return e;
    }
    if (m.n_long > 0) {
        hipLaunchKernelGGL(dev::csr_long_finalize_kernel, dim3((m.n_long + 255) / 256), dim3(256), 0, st, m.long_rows,
                           m.long_first, m.carry, m.n_long, y);
    }
    return (int)hipGetLastError();
}

int launch_csr_sliced(const DevCsrSliced& m, const float* x, float* y, hipStream_t st) {
    for (int s = 0; s < m.nslices; ++s) {
        const int b0 = m.slice_block_begin[s], nb = m.slice_block_begin[s + 1] - b0;
        if (nb == 0) continue;
        hipLaunchKernelGGL(dev::csr_slice_kernel, dim3((nb + 3) / 4), dim3(256), 0, st, m.counts + (size_t)s * m.rows,
                           m.col, m.val, m.blocks + b0, nb, x, y, m.carry);
        int e = (int)hipGetLastError();
        if (e) return e;
    }
    if (m.n_long > 0) {
        hipLaunchKernelGGL(dev::csr_long_finalize_kernel, dim3((m.n_long + 255) / 256), dim3(256), 0, st, m.long_rows,
                           m.long_first, m.carry, m.n_long, y);
    }
    return (int)hipGetLastError();
}

}  // namespace clspmv
