// Division and remainder by a run-time divisor d >= 1 without a divide
// instruction: the round-up multiply-high method of Granlund and Montgomery
// ("Division by Invariant Integers using Multiplication", PLDI 1994, fig. 4.1),
// exact for every 32-bit dividend.  With l = ceil(log2 d),
//   m = floor(2^32 (2^l - d) / d) + 1,   t = mulhi(m, x),
//   x / d = (t + ((x - t) >> min(l, 1))) >> max(l - 1, 0).
// On the device the product is one v_mul_hi_u32: a remainder costs about six
// instructions instead of the thirty or so of a generic 32-bit `%`.
// Plain C++ (constexpr functions are callable from HIP device code), so the
// host tests check it against `%` directly.
#pragma once

#include <stdint.h>

namespace fls {

struct FastDiv {
    uint32_t d, m, sh1, sh2;
};

constexpr FastDiv make_fastdiv(uint32_t d) {
    uint32_t l = 0;
    while (l < 32 && (uint64_t{1} << l) < d) ++l;
    const uint64_t num = ((uint64_t{1} << l) - d) << 32;  // 2^l - d < d <= 2^32: fits
    return FastDiv{d, (uint32_t)(num / (d ? d : 1) + 1), l > 0 ? 1u : 0u, l > 0 ? l - 1 : 0u};
}

constexpr uint32_t fast_div(uint32_t x, FastDiv f) {
    const uint32_t t = (uint32_t)(((uint64_t)x * f.m) >> 32);
    return (t + ((x - t) >> f.sh1)) >> f.sh2;
}

constexpr uint32_t fast_mod(uint32_t x, FastDiv f) { return x - fast_div(x, f) * f.d; }

}  // namespace fls
