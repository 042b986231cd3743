// Device helpers shared by the SpMV kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace clspmv {
namespace dev {

constexpr int kWave = 64;

// Streaming loads for matrix arrays that are read exactly once per SpMV: the
// nontemporal hint keeps them from evicting the reused x vector from L2.
template <class T>
__device__ __forceinline__ T ld_stream(const T* p) {
    return __builtin_nontemporal_load(p);
}
typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ int4 ld_stream(const int4* p) {
    i32x4 v = __builtin_nontemporal_load(reinterpret_cast<const i32x4*>(p));
    return make_int4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ float4 ld_stream(const float4* p) {
    f32x4 v = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(p));
    return make_float4(v.x, v.y, v.z, v.w);
}

// Blocks per step of the sliced BSR kernels (FP32 and mixed precision):
// the column, x and value loads of that many blocks are issued together.
constexpr int kBsrGroupSmall = 4;  // blocks of <= 16 values
__host__ __device__ constexpr int bsr_group(int values_per_block) {
    return values_per_block <= 16 ? kBsrGroupSmall : 2;
}

// x[c .. c+BW-1] for a block starting at column c, zero at and past `cols`
// (requires cols >= BW): BW consecutive floats are read from
// min(c, cols - BW) -- one multi-dword load, no branch -- and shifted into
// place; the shift is nonzero only for a block overlapping the end of x.
template <int BW>
__device__ __forceinline__ void load_xblock(const float* __restrict__ x, int c, int cols, float (&xv)[BW]) {
    const int cb = min(c, cols - BW);
    float v[BW];
#pragma unroll
    for (int w = 0; w < BW; ++w) v[w] = x[cb + w];
    const int d = c - cb;
#pragma unroll
    for (int w = 0; w < BW; ++w) {
        float o = 0.f;
#pragma unroll
        for (int j = 0; j < BW; ++j)
            if (w + j < BW) o = (d == j) ? v[w + j] : o;
        xv[w] = o;
    }
}

// Stages x[w0 .. w0 + n) (zero at and past `cols`) into xs[0 .. n) (xs
// 16-byte aligned) with the whole workgroup (T threads): 16-byte loads of x
// at 4-byte alignment, up to 10 per thread in flight (a 40960-float window
// in one round); requires cols >= 4.  The caller adds the barrier.
typedef float f32x4a4 __attribute__((ext_vector_type(4), aligned(4)));
template <int T>
__device__ __forceinline__ void stage_xwindow(const float* __restrict__ x, int w0, int n, int cols, float* xs) {
    constexpr int U = 10;
    // loads start at most at lim - 4: past the window (or x) they re-read
    // its last 16 bytes, a cache hit, rather than x beyond it
    const int lim = min(w0 + max(n, 4), cols);
    for (int i0 = 0; i0 < n; i0 += 4 * U * T) {
        f32x4a4 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int c = w0 + i0 + 4 * (u * T + (int)threadIdx.x);
            if (i0 + 4 * u * T < n) v[u] = *reinterpret_cast<const f32x4a4*>(x + min(c, lim - 4));  // uniform
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = i0 + 4 * (u * T + (int)threadIdx.x);
            const int c = w0 + i;
            if (i + 4 <= n && c + 4 <= cols) {
                *reinterpret_cast<float4*>(xs + i) = make_float4(v[u].x, v[u].y, v[u].z, v[u].w);
            } else if (i < n) {  // the end of the window or of x: element by element
                const int d = c - min(c, lim - 4);  // shift of the clamped load
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    float e = 0.f;
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        if (q == d + j) e = v[u][q];
                    if (i + j < n) xs[i + j] = c + j < cols ? e : 0.f;
                }
            }
        }
    }
}

// Sum over all 64 lanes; every lane receives the total.
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, kWave);
    return v;
}

// Sum over aligned groups of L lanes (L a power of two <= 64).
template <int L>
__device__ __forceinline__ float group_sum(float v) {
#pragma unroll
    for (int m = L / 2; m >= 1; m >>= 1) v += __shfl_xor(v, m, kWave);
    return v;
}

// Orders LDS traffic among the lanes of one wavefront (no workgroup barrier:
// the waves of a workgroup may be on different code paths).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X.  This
// maps hardware block ids so that each XCD processes one contiguous range of
// logical tiles, keeping the x windows that neighbouring tiles share inside
// one XCD's L2.  Bijective for any n.
__device__ __forceinline__ int xcd_tile(int b, int n) {
    const int q = n >> 3, rem = n & 7;
    const int x = b & 7, idx = b >> 3;
    return (x < rem) ? x * (q + 1) + idx : rem * (q + 1) + (x - rem) * q + idx;
}

}  // namespace dev
}  // namespace clspmv
