// Device decoders of the mixed-precision value codecs (host encoders:
// host/codec.hpp).  Products and sums stay FP32.
#pragma once
#include <hip/hip_fp16.h>
#include <stdint.h>

namespace clspmv {
namespace dev {

// BF16 hi + lo split in one word: hi in the upper half, lo in the lower.
__device__ __forceinline__ float bf16x2_value(uint32_t u) {
    return __uint_as_float(u & 0xffff0000u) + __uint_as_float(u << 16);
}
__device__ __forceinline__ float f16_value(uint16_t h) { return __half2float(__ushort_as_half(h)); }
// Signed 16-bit fixed point: the low (lo) or high (hi) half of a word,
// sign-extended and converted (exact: |q| <= 32767).
__device__ __forceinline__ float i16_lo(uint32_t w) { return (float)((int)(w << 16) >> 16); }
__device__ __forceinline__ float i16_hi(uint32_t w) { return (float)((int)w >> 16); }
// E4M3: normal (e >= 1) 2^(e-7) * (1 + m/8), subnormal (e = 0) m * 2^-9;
// the caller's scale absorbs the magnitude.  Branch-free: the 7 magnitude
// bits shifted into place form the FP32 2^(e-127) * (1 + m/8) for e >= 1 and
// the FP32 subnormal m/8 * 2^-126 for e = 0, so one multiply by 2^120 gives
// both (the kernels keep FP32 denormals: denorm mode 3 in their
// descriptors); the sign bit is copied over.
__device__ __forceinline__ float e4m3_value(uint32_t q) {
    const float mag = __uint_as_float((q & 0x7fu) << 20) * 0x1p120f;
    return __uint_as_float(__float_as_uint(mag) | ((q & 0x80u) << 24));
}

}  // namespace dev
}  // namespace clspmv
