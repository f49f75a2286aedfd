// kd_expf: single-precision exp used by the render path on host and device,
// so the device kernel and the host oracle produce identical bits.
//
// Method: Cody & Waite (1980) argument reduction x = k*ln2 + r with ln2
// split into a 16-bit head (k*head is exact for |k| <= 255) and a tail, then
// the degree-7 Taylor polynomial of e^r on |r| <= ln2/2 (truncation error
// < 1e-8 relative) in Horner form, then scaling by 2^k built from the
// exponent bits.  Rounding of k uses the 1.5*2^23 add/subtract trick
// (round-half-even, valid for |x*log2e| < 2^22).  Every step is a single
// IEEE-754 binary32 add or multiply; the translation unit must be compiled
// with -ffp-contract=off on both sides.  Accuracy against a correctly
// rounded exp is checked by the test suite (max error stated there).
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define KD_HD __host__ __device__
#else
#define KD_HD
#endif

KD_HD static inline float kd_bits_to_float(uint32_t u) {
    union {
        uint32_t u;
        float f;
    } c;
    c.u = u;
    return c.f;
}

KD_HD static inline float kd_expf(float x) {
    if (!(x == x)) return x;                                   // NaN
    if (x > 88.72283935546875f) return kd_bits_to_float(0x7F800000u); // overflow -> +inf
    if (x < -103.972084045410156f) return 0.0f;                // underflow -> +0
    const float kLog2e = 1.44269502162933349609375f;           // nearest float to 1/ln2
    const float kLn2Hi = 0.693145751953125f;                   // 0x3F317200
    const float kLn2Lo = 1.428606765330187045e-06f;            // 0x35BFBE8E
    const float kRound = 12582912.0f;                          // 1.5 * 2^23
    float kf = x * kLog2e;
    kf = kf + kRound;
    kf = kf - kRound;
    float r = x - kf * kLn2Hi;
    r = r - kf * kLn2Lo;
    float p = 1.98412698412698413e-4f; // 1/7!
    p = p * r + 1.38888888888888889e-3f; // 1/6!
    p = p * r + 8.33333333333333333e-3f; // 1/5!
    p = p * r + 4.16666666666666667e-2f; // 1/4!
    p = p * r + 1.66666666666666667e-1f; // 1/3!
    p = p * r + 0.5f;
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    int k = (int)kf;
    if (k > 127) {
        p = p * kd_bits_to_float(0x7F000000u); // 2^127
        k -= 127;
    } else if (k < -126) {
        p = p * kd_bits_to_float((uint32_t)(k + 64 + 127) << 23); // 2^(k+64), normal
        return p * kd_bits_to_float(0x1F800000u);                   // 2^-64
    }
    return p * kd_bits_to_float((uint32_t)(k + 127) << 23);
}
