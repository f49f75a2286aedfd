// Exhaustive check of kdgrt::f32_to_f16_up (csrc/f16_up.hpp): for every
// positive finite float up to 65504 the result must be the bit pattern of
// the smallest half-precision value >= x, found here by a binary search in
// the table of all finite non-negative halves (whose bit patterns ascend
// with their values); plus the special inputs.  Prints the mismatch count.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <vector>

#include "f16_up.hpp"

int main() {
    std::vector<float> half(0x7C00);
    for (uint32_t b = 0; b < 0x7C00u; ++b) {
        const int ex = (int)(b >> 10);
        const uint32_t fr = b & 0x3FFu;
        half[b] = ex == 0 ? std::ldexp((float)fr, -24) : std::ldexp((float)(1024u + fr), ex - 25);
    }
    long bad = 0;
    for (uint32_t b = 1; b < 0x7C00u; ++b) bad += half[b] > half[b - 1] ? 0 : 1; // table ascends
    float top_f = 65504.0f;
    uint32_t top;
    std::memcpy(&top, &top_f, 4);
#pragma omp parallel for reduction(+ : bad) schedule(static, 1 << 20)
    for (int64_t i = 1; i <= (int64_t)top; ++i) {
        const uint32_t u = (uint32_t)i;
        float x;
        std::memcpy(&x, &u, 4);
        const uint32_t want = (uint32_t)(std::lower_bound(half.begin(), half.end(), x) - half.begin());
        if (kdgrt::f32_to_f16_up(x) != want) ++bad;
    }
    struct Case {
        float x;
        uint32_t want;
    } cases[] = {{0.0f, 0u},
                 {-0.0f, 0u},
                 {-1.0f, 0u},
                 {65504.0f, 0x7BFFu},
                 {std::nextafter(65504.0f, 1e30f), 0x7C00u},
                 {std::numeric_limits<float>::infinity(), 0x7C00u},
                 {std::numeric_limits<float>::quiet_NaN(), 0x7C00u},
                 {std::numeric_limits<float>::denorm_min(), 1u},
                 {0.0010987452f, 0x1481u}};
    for (const Case& c : cases)
        if (kdgrt::f32_to_f16_up(c.x) != c.want) {
            std::printf("special %.9g: got 0x%04x want 0x%04x\n", c.x, kdgrt::f32_to_f16_up(c.x), c.want);
            ++bad;
        }
    std::printf("floats checked %u, mismatches %ld\n", top, bad);
    return bad != 0;
}
