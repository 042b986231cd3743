// Value encodings of the mixed-precision parts (host side): FP32 -> BF16
// (round to nearest even), FP32 -> FP16 (the compiler's _Float16 conversion,
// round to nearest even), FP32 -> FP8 E4M3 (OCP 8-bit format: 4 exponent
// bits, bias 7, 3 mantissa bits, largest finite 448, no infinities; round to
// nearest even, saturating).  The device decoders are in kernels/k_mixed.hip.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>

namespace clspmv {

inline uint16_t f32_to_bf16(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    const uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    return (uint16_t)(u >> 16);
}
inline float bf16_to_f32(uint16_t h) {
    const uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

inline uint16_t f32_to_f16(float f) {
    const _Float16 h = (_Float16)f;
    uint16_t u;
    std::memcpy(&u, &h, 2);
    return u;
}

// E4M3: value = (-1)^s * 2^(e-7) * (1 + m/8) for e >= 1, 2^-6 * m/8 for e = 0.
inline float e4m3_to_f32(uint8_t q) {
    const int s = q >> 7, e = (q >> 3) & 15, m = q & 7;
    const float v = e ? std::ldexp(1.0f + m / 8.0f, e - 7) : std::ldexp(m / 8.0f, -6);
    return s ? -v : v;
}
inline uint8_t f32_to_e4m3(float f) {
    // codes 0x00..0x7e of one sign are increasing in value: binary search,
    // then the nearer neighbour (ties to the even code)
    static const struct Table {
        float v[127];
        Table() {
            for (int q = 0; q < 127; ++q) v[q] = e4m3_to_f32((uint8_t)q);
        }
    } T;
    const uint8_t sign = std::signbit(f) ? 0x80 : 0x00;
    const float a = std::fabs(f);
    if (!(a == a)) return sign | 0x7f;  // NaN
    if (a >= T.v[126]) return sign | 0x7e;  // saturate to the largest finite
    int lo = 0, hi = 126;  // v[lo] <= a < v[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (T.v[mid] <= a)
            lo = mid;
        else
            hi = mid;
    }
    const float dl = a - T.v[lo], dh = T.v[hi] - a;
    const int q = dl < dh ? lo : dh < dl ? hi : ((lo & 1) ? hi : lo);
    return (uint8_t)(sign | q);
}

}  // namespace clspmv
