// Batched multi-right-hand-side SpMV: Y_i += A_i X_i for many CSR matrices
// in one persistent launch, each with k_i in {1, 4, 16} right-hand sides.
//
// X_i is stored row-major (x[col*k + j]) and Y_i likewise, so the k values a
// nonzero multiplies are contiguous: one gather of a 64-byte line serves all
// 16 products of a k = 16 matrix, where k single-vector SpMVs would gather k
// times.
//
// Work items (planned on the host, any mix of matrices):
//   STREAM {mat, r0, r1, 0}: rows [r0, r1) with at most kSpmmStageNnz
//     entries.  The wave copies the entries' columns and values into its LDS
//     with coalesced loads, then "slots" of lanes each own rows: k/4 lanes
//     (one for k <= 4) hold a row's float4 (or float) pieces of x[col][:],
//     times as many lanes as the item's row count leaves free, which stride
//     the row's entries and are summed with shuffles; the sum is added to
//     y[row][:].
//   LONG {mat, row, -(slot+1), chunk}: kSpmmLongChunk entries of a long row;
//     every slot strides the chunk, the slots' partials are summed with
//     shuffles and written to carry slot `slot` (16 floats); spmm_finalize
//     adds the carries of each long row in chunk order.
//   BLOCK {mat, b0, b1, 0} of a BSR matrix (dense bdim x bdim blocks,
//     bdim = 2..4, k = 4 or 16): block rows [b0, b1), at most 16 / (k/4),
//     on the matrix cores.  One v_mfma_f32_4x4x1f32 is 16 independent 4x4
//     outer products; here each is (block row g, right-hand-side group q):
//     rows of the block row's current block, column c, times x[col + c]
//     [4q .. 4q+3], so bdim instructions advance every block row of the
//     item by one block (see spmm_block).
// Waves take items i, i + W, i + 2W, ... (W = waves in the grid), so the
// launch is one grid whatever the number of matrices.  No atomics: results
// are deterministic.
#include "kcommon.hpp"
#include "launch.hpp"

namespace clspmv {
namespace dev {

template <int K>
struct SpmmLanes {
    static constexpr int W = K >= 4 ? 4 : 1;       // floats per lane
    static constexpr int V = K >= 4 ? K / 4 : 1;   // lanes per right-hand-side vector
    static constexpr int SLOTS = 64 / V;
};

template <int W>
struct Vec;
template <>
struct Vec<1> {
    float a;
    __device__ void zero() { a = 0.f; }
    __device__ void fma_from(const float* p, float f) { a += f * p[0]; }
    __device__ static Vec load(const float* p) { return Vec{p[0]}; }
    __device__ void fma(const Vec& x, float f) { a += f * x.a; }
    __device__ void add_to(float* p) const { p[0] += a; }
    __device__ void store(float* p) const { p[0] = a; }
    __device__ void shfl_add(int m) { a += __shfl_xor(a, m, 64); }
};
template <>
struct Vec<4> {
    float4 a;
    __device__ void zero() { a = make_float4(0.f, 0.f, 0.f, 0.f); }
    __device__ void fma_from(const float* p, float f) { fma(load(p), f); }
    __device__ static Vec load(const float* p) { return Vec{*reinterpret_cast<const float4*>(p)}; }
    __device__ void fma(const Vec& x, float f) {
        a.x += f * x.a.x;
        a.y += f * x.a.y;
        a.z += f * x.a.z;
        a.w += f * x.a.w;
    }
    __device__ void add_to(float* p) const {
        float4 y = *reinterpret_cast<float4*>(p);
        y.x += a.x;
        y.y += a.y;
        y.z += a.z;
        y.w += a.w;
        *reinterpret_cast<float4*>(p) = y;
    }
    __device__ void store(float* p) const { *reinterpret_cast<float4*>(p) = a; }
    __device__ void shfl_add(int m) {
        a.x += __shfl_xor(a.x, m, 64);
        a.y += __shfl_xor(a.y, m, 64);
        a.z += __shfl_xor(a.z, m, 64);
        a.w += __shfl_xor(a.w, m, 64);
    }
};

template <int K>
__device__ void spmm_stream(const SpmmMat& m, int r0, int r1, int* s_col, float* s_val, int lane) {
    using Ln = SpmmLanes<K>;
    const int n0 = m.row_ptr[r0], n = m.row_ptr[r1] - n0;
    for (int i = lane; i < n; i += 64) {
        s_col[i] = m.col[n0 + i];
        s_val[i] = m.val[n0 + i];
    }
    wave_lds_sync();
    // lanes per row: V (one right-hand-side vector) times `split` lanes that
    // stride the row's entries, as many as the item's rows leave free
    const int nrows = r1 - r0;
    int split = 1;
    while (split * 2 * Ln::V * nrows <= 64) split *= 2;
    const int lpr = Ln::V * split, slots = 64 / lpr;
    const int slot = lane / lpr, sp = (lane / Ln::V) % split, v = lane % Ln::V;
    const float* xb = m.x + v * Ln::W;
    for (int r = r0 + slot; r < r1; r += slots) {
        const int a = m.row_ptr[r] - n0, b = m.row_ptr[r + 1] - n0;
        Vec<Ln::W> acc;
        acc.zero();
        int j = a + sp;
        // four independent gathers in flight per lane, then their products
        for (; j + 3 * split < b; j += 4 * split) {
            const Vec<Ln::W> x0 = Vec<Ln::W>::load(xb + (size_t)s_col[j] * K);
            const Vec<Ln::W> x1 = Vec<Ln::W>::load(xb + (size_t)s_col[j + split] * K);
            const Vec<Ln::W> x2 = Vec<Ln::W>::load(xb + (size_t)s_col[j + 2 * split] * K);
            const Vec<Ln::W> x3 = Vec<Ln::W>::load(xb + (size_t)s_col[j + 3 * split] * K);
            acc.fma(x0, s_val[j]);
            acc.fma(x1, s_val[j + split]);
            acc.fma(x2, s_val[j + 2 * split]);
            acc.fma(x3, s_val[j + 3 * split]);
        }
        for (; j < b; j += split) acc.fma_from(xb + (size_t)s_col[j] * K, s_val[j]);
        // the lanes of one row run the same iterations, so the shuffles below
        // only read lanes that are active
        for (int msk = Ln::V; msk < lpr; msk <<= 1) acc.shfl_add(msk);
        if (sp == 0) acc.add_to(m.y + (size_t)r * K + v * Ln::W);
    }
    wave_lds_sync();  // the next item reuses this wave's LDS
}

template <int K>
__device__ void spmm_long(const SpmmMat& m, int row, int slot_id, int chunk, float* carry, int lane) {
    using Ln = SpmmLanes<K>;
    const int s = lane / Ln::V, v = lane % Ln::V;
    const int n0 = m.row_ptr[row] + chunk * kSpmmLongChunk;
    const int n1 = min(m.row_ptr[row + 1], n0 + kSpmmLongChunk);
    Vec<Ln::W> acc;
    acc.zero();
    const float* xb = m.x + v * Ln::W;
    int j = n0 + s;
    for (; j + 3 * Ln::SLOTS < n1; j += 4 * Ln::SLOTS) {
        const int c0 = m.col[j], c1 = m.col[j + Ln::SLOTS], c2 = m.col[j + 2 * Ln::SLOTS], c3 = m.col[j + 3 * Ln::SLOTS];
        const Vec<Ln::W> x0 = Vec<Ln::W>::load(xb + (size_t)c0 * K);
        const Vec<Ln::W> x1 = Vec<Ln::W>::load(xb + (size_t)c1 * K);
        const Vec<Ln::W> x2 = Vec<Ln::W>::load(xb + (size_t)c2 * K);
        const Vec<Ln::W> x3 = Vec<Ln::W>::load(xb + (size_t)c3 * K);
        acc.fma(x0, m.val[j]);
        acc.fma(x1, m.val[j + Ln::SLOTS]);
        acc.fma(x2, m.val[j + 2 * Ln::SLOTS]);
        acc.fma(x3, m.val[j + 3 * Ln::SLOTS]);
    }
    for (; j < n1; j += Ln::SLOTS) acc.fma_from(xb + (size_t)m.col[j] * K, m.val[j]);
#pragma unroll
    for (int msk = Ln::V; msk < 64; msk <<= 1) acc.shfl_add(msk);
    if (s == 0) acc.store(carry + (size_t)slot_id * 16 + v * Ln::W);
}

typedef float f32x4 __attribute__((ext_vector_type(4)));

// One BLOCK item: block rows [b0, b1) of a BSR matrix with B x B blocks and
// K in {4, 16} right-hand sides.  v_mfma_f32_4x4x1f32 works on 16 blocks of
// four lanes: lane 4 * blk + t gives row t of the block's 4 x 1 operand A
// and column t of its 1 x 4 operand B, and holds column t of the 4 x 4
// result (rows in its four accumulator registers).  Block blk is (block row
// g = blk / Q, right-hand-side group q = blk % Q), Q = K / 4; so with
// A = column c of g's current block (rows >= B read 0) and B = x[col + c]
// [4q .. 4q+3], the c = 0 .. B-1 steps add that block's product to the
// block row's 4 x 4 tile of Y.  Every block row of the item advances by one
// block per B steps; rows with fewer blocks feed zeros (and a valid x).
template <int K, int B>
__device__ void spmm_block(const SpmmMat& m, int b0, int b1, int lane) {
    constexpr int Q = K / 4, G = 16 / Q;
    const int blk = lane >> 2, t = lane & 3;
    const int g = blk / Q, q = blk % Q;
    const int br = b0 + g;
    const bool row_ok = br < b1;
    const int s = row_ok ? m.row_ptr[br] : 0, e = row_ok ? m.row_ptr[br + 1] : 0;
    // blocks of the longest block row of the item (wave-uniform)
    int n = e - s;
#pragma unroll
    for (int d = 4; d < 64; d <<= 1) n = max(n, __shfl_xor(n, d, 64));
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    const float* xq = m.x + q * 4 + t;
    for (int j = 0; j < n; ++j) {
        const bool on = j < e - s;
        const int kb = on ? s + j : 0;
        const int c0 = on ? m.col[kb] : 0;
        const float* a = m.val + (size_t)kb * (B * B) + t * B;
        float av[B], xv[B];
#pragma unroll
        for (int c = 0; c < B; ++c) {
            av[c] = (on && t < B) ? a[c] : 0.f;
            const int col = min(c0 + c, m.cols - 1);  // block columns past the last read zeros in A
            xv[c] = xq[(size_t)col * K];
        }
#pragma unroll
        for (int c = 0; c < B; ++c) acc = __builtin_amdgcn_mfma_f32_4x4x1f32(av[c], xv[c], acc, 0, 0, 0);
    }
    if (row_ok) {
#pragma unroll
        for (int v = 0; v < B; ++v) {
            const int r = br * B + v;
            if (r < m.rows) m.y[(size_t)r * K + q * 4 + t] += acc[v];
        }
    }
    (void)G;
}

__global__ __launch_bounds__(256) void spmm_kernel(const SpmmMat* __restrict__ mats, const int4* __restrict__ items,
                                                   int n_items, float* __restrict__ carry) {
    __shared__ int s_col[4][kSpmmStageNnz];
    __shared__ float s_val[4][kSpmmStageNnz];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int waves = gridDim.x * 4;
    for (int it = blockIdx.x * 4 + w; it < n_items; it += waves) {
        const int4 d = items[it];
        const SpmmMat m = mats[d.x];
        if (m.bdim) {
            // BSR items (host-built: k = 4 or 16, bdim = 2..4)
            switch (m.k * 10 + m.bdim) {
                case 42: spmm_block<4, 2>(m, d.y, d.z, lane); break;
                case 43: spmm_block<4, 3>(m, d.y, d.z, lane); break;
                case 44: spmm_block<4, 4>(m, d.y, d.z, lane); break;
                case 162: spmm_block<16, 2>(m, d.y, d.z, lane); break;
                case 163: spmm_block<16, 3>(m, d.y, d.z, lane); break;
                default: spmm_block<16, 4>(m, d.y, d.z, lane); break;
            }
        } else if (d.z >= 0) {
            switch (m.k) {
                case 1: spmm_stream<1>(m, d.y, d.z, s_col[w], s_val[w], lane); break;
                case 4: spmm_stream<4>(m, d.y, d.z, s_col[w], s_val[w], lane); break;
                default: spmm_stream<16>(m, d.y, d.z, s_col[w], s_val[w], lane); break;
            }
        } else {
            const int slot = -d.z - 1;
            switch (m.k) {
                case 1: spmm_long<1>(m, d.y, slot, d.w, carry, lane); break;
                case 4: spmm_long<4>(m, d.y, slot, d.w, carry, lane); break;
                default: spmm_long<16>(m, d.y, slot, d.w, carry, lane); break;
            }
        }
    }
}

// One thread per (long row, right-hand side): y += carries in chunk order,
// summed in FP64 (the cross-workgroup part of a long row's sum).
__global__ __launch_bounds__(256) void spmm_finalize_kernel(const SpmmMat* __restrict__ mats,
                                                            const int4* __restrict__ longs, int n_long,
                                                            const float* __restrict__ carry) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int i = t >> 4, j = t & 15;
    if (i >= n_long) return;
    const int4 d = longs[i];  // {mat, row, first slot, chunks}
    const SpmmMat m = mats[d.x];
    if (j >= m.k) return;
    double s = 0.0;
    for (int c = 0; c < d.w; ++c) s += (double)carry[(size_t)(d.z + c) * 16 + j];
    float* yp = m.y + (size_t)d.y * m.k + j;
    *yp = (float)((double)*yp + s);
}

}  // namespace dev

int launch_spmm(const DevSpmm& b, hipStream_t st) {
    if (b.n_items > 0) {
        hipLaunchKernelGGL(dev::spmm_kernel, dim3(b.grid), dim3(256), 0, st, b.mats, b.items, b.n_items, b.carry);
        int e = (int)hipGetLastError();
        if (e) return e;
    }
    if (b.n_long > 0) {
        hipLaunchKernelGGL(dev::spmm_finalize_kernel, dim3((b.n_long * 16 + 255) / 256), dim3(256), 0, st, b.mats,
                           b.longs, b.n_long, b.carry);
        return (int)hipGetLastError();
    }
    return 0;
}

}  // namespace clspmv
