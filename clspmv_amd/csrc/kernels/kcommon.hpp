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
