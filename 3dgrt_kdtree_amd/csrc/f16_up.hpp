// Host-side rounding of a non-negative float up to IEEE half precision
// (the bound-sphere records store their radius and dilation rounded up).
// Plain C++ so that a CPU test checks it against every float.
#pragma once
#include <cstdint>
#include <cstring>

namespace kdgrt {

inline uint16_t f32_to_f16_up(float x) {
    // smallest half >= x for x >= 0 (inf and NaN map to +inf), from the
    // bits: the host forms of hip_fp16's directed conversions
    // (__float2half_ru) round to nearest, which can land below x.
    if (!(x <= 65504.0f)) return 0x7C00u;
    if (x <= 0.0f) return 0;
    uint32_t u;
    std::memcpy(&u, &x, 4);
    const int e = (int)((u >> 23) & 0xFFu) - 127;       // x = 1.f * 2^e (f32 subnormals: e = -127)
    const uint32_t m = (u & 0x7FFFFFu) | 0x800000u;      // 24-bit significand
    if (e >= -14) {
        // normal half: keep 10 fraction bits, round the 13 dropped up; a
        // carry out of the fraction moves to the next exponent, the next
        // half up (x <= 65504 keeps the result finite)
        uint32_t b = ((uint32_t)(e + 15) << 10) | ((m >> 13) & 0x3FFu);
        if (m & 0x1FFFu) ++b;
        return (uint16_t)b;
    }
    // subnormal half: k units of 2^-24 with k = ceil(x * 2^24) = ceil(m * 2^(e+1))
    const int s = -(e + 1); // >= 14
    const uint32_t k = s >= 32 ? 0u : (m >> s);
    const bool rest = s >= 32 || (m & ((1u << s) - 1u)) != 0;
    return (uint16_t)(k + (rest ? 1u : 0u));
}

} // namespace kdgrt
