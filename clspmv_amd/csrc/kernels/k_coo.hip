// COO SpMV (y += A*x) as a segmented reduction over the row-sorted entries.
//
// The reference splits the entries into per-workgroup ranges, runs an LDS
// segmented scan with a barrier per step and folds the per-group carries in a
// second one-workgroup kernel (A/kernels/spmv_coo.cl:4-200).  Here one
// wavefront owns a 512-entry chunk and each lane owns 8 consecutive entries,
// loaded as two dwordx4 per array (a wave's loads are contiguous):
//   1. lane-local pass: segments that start and end inside the lane are
//      complete and written to y directly (rows are unique to a lane there);
//   2. a 6-step shuffle segmented scan carries lane tails into the head
//      segment of the next lane(s); completed heads are written to y;
//   3. the chunk's first row (may continue from the previous chunk) and last
//      row (may continue into the next) go to a per-chunk carry pair in FP64;
//   4. coo_carry_kernel adds each run of equal-row carries in order.
// No atomics; results are deterministic.  Entries are padded to a multiple of
// 512 with (last row, last col, 0.0), as the reference pads to 32; arrays
// that end earlier (nnz below the chunk multiple) are padded the same way
// as the last chunk is loaded.
#include "kcommon.hpp"
#include "launch.hpp"

namespace clspmv {
namespace dev {

__global__ __launch_bounds__(256) void coo_segment_kernel(const int* __restrict__ rowi, const int* __restrict__ coli,
                                                          const float* __restrict__ vali, int nnz, int chunks,
                                                          const float* __restrict__ x, float* __restrict__ y,
                                                          int* __restrict__ carry_row,
                                                          double* __restrict__ carry_val) {
    const int lane = threadIdx.x & 63;
    const int chunk = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (chunk >= chunks) return;
    const size_t base = (size_t)chunk * kCooChunk + (size_t)lane * kCooPerLane;
    int r[8], c[8];
    float v[8];
    if (base + kCooPerLane > (size_t)nnz) {
        // past the end of unpadded arrays: (last row, last col, 0.0), every
        // load at a valid index
        const int last = nnz - 1;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const bool in = base + (size_t)i < (size_t)nnz;
            const size_t k = in ? base + (size_t)i : (size_t)last;
            r[i] = rowi[k];
            c[i] = coli[k];
            v[i] = in ? vali[k] : 0.f;
        }
    } else {
        const int4* rp4 = reinterpret_cast<const int4*>(rowi + base);
        const int4* cp4 = reinterpret_cast<const int4*>(coli + base);
        const float4* vp4 = reinterpret_cast<const float4*>(vali + base);
        int4 ra = ld_stream(rp4), rb = ld_stream(rp4 + 1);
        int4 ca = ld_stream(cp4), cb = ld_stream(cp4 + 1);
        float4 va = ld_stream(vp4), vb = ld_stream(vp4 + 1);
        r[0] = ra.x; r[1] = ra.y; r[2] = ra.z; r[3] = ra.w; r[4] = rb.x; r[5] = rb.y; r[6] = rb.z; r[7] = rb.w;
        c[0] = ca.x; c[1] = ca.y; c[2] = ca.z; c[3] = ca.w; c[4] = cb.x; c[5] = cb.y; c[6] = cb.z; c[7] = cb.w;
        v[0] = va.x; v[1] = va.y; v[2] = va.z; v[3] = va.w; v[4] = vb.x; v[5] = vb.y; v[6] = vb.z; v[7] = vb.w;
    }
    float p[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) p[i] = v[i] * x[c[i]];

    // 1. lane-local segments.
    const int head_row = r[0];
    int cur = r[0];
    float acc = 0.f, head_sum = 0.f;
    bool brk = false;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        if (r[i] != cur) {
            if (!brk) {
                head_sum = acc;
                brk = true;
            } else {
                y[cur] += acc;
            }
            cur = r[i];
            acc = 0.f;
        }
        acc += p[i];
    }
    const int tail_row = cur;
    const float tail_sum = acc;

    // 2. segmented inclusive scan of the lane tails across the wave.  A lane
    //    starts a new segment if it contains a row break or its row differs
    //    from the previous lane's tail row.
    const int prev_tail_row = __shfl_up(tail_row, 1, 64);
    int flag = (lane == 0) || brk || (prev_tail_row != tail_row);
    float cval = tail_sum;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        float up_v = __shfl_up(cval, d, 64);
        int up_f = __shfl_up(flag, d, 64);
        if (lane >= d && !flag) {
            cval += up_v;
            flag = up_f;
        }
    }
    // cval = tail of this lane including carries from earlier lanes.
    const float prev_c = __shfl_up(cval, 1, 64);
    const int chunk_first_row = __shfl(head_row, 0, 64);
    // A segment that ends exactly at the boundary before this lane is
    // complete: (prev_tail_row, prev_c).
    const bool boundary_end = lane > 0 && prev_tail_row != head_row;
    if (boundary_end) {
        if (prev_tail_row == chunk_first_row) {
            carry_row[2 * chunk] = prev_tail_row;
            carry_val[2 * chunk] = (double)prev_c;
        } else {
            y[prev_tail_row] += prev_c;
        }
    }
    // A row break inside the lane completes the head segment.
    if (brk) {
        const float h = head_sum + ((lane > 0 && !boundary_end) ? prev_c : 0.f);
        if (head_row == chunk_first_row) {
            carry_row[2 * chunk] = head_row;
            carry_val[2 * chunk] = (double)h;
        } else {
            y[head_row] += h;
        }
    }
    if (lane == 63) {
        // The whole chunk is one row: its first-row carry is empty.
        if (tail_row == chunk_first_row) {
            carry_row[2 * chunk] = chunk_first_row;
            carry_val[2 * chunk] = 0.0;
        }
        carry_row[2 * chunk + 1] = tail_row;
        carry_val[2 * chunk + 1] = (double)cval;
    }
}

// Each run of equal rows in the (row-sorted) carry list is summed in FP64 by
// the lane that owns its first entry.
__global__ __launch_bounds__(256) void coo_carry_kernel(const int* __restrict__ carry_row,
                                                        const double* __restrict__ carry_val, int n,
                                                        float* __restrict__ y, int rows) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int r = carry_row[i];
    if (r < 0 || r >= rows) return;  // defensive: every slot is written by coo_segment_kernel
    if (i > 0 && carry_row[i - 1] == r) return;
    double s = 0.0;
    for (int k = i; k < n && carry_row[k] == r; ++k) s += carry_val[k];
    y[r] = (float)((double)y[r] + s);
}

}  // namespace dev

int launch_coo(const DevCoo& m, const float* x, float* y, hipStream_t st) {
    if (m.chunks == 0) return 0;
    hipLaunchKernelGGL(dev::coo_segment_kernel, dim3((m.chunks + 3) / 4), dim3(256), 0, st, m.row, m.col, m.val,
                       m.nnz, m.chunks, x, y, m.carry_row, m.carry_val);
    int e = (int)hipGetLastError();
    if (e) return e;
    const int n = 2 * m.chunks;
    hipLaunchKernelGGL(dev::coo_carry_kernel, dim3((n + 255) / 256), dim3(256), 0, st, m.carry_row, m.carry_val, n, y,
                       m.rows);
    return (int)hipGetLastError();
}

}  // namespace clspmv
