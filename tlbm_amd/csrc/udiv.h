// udiv.h -- division by a launch-time constant without an integer divide.
//
// For a divisor 2 <= d < 2^31 that is not a power of two, with
// l = ceil(log2 d) (so 2^(l-1) < d < 2^l) and m = ceil(2^(31+l) / d)
// (2^31 <= m < 2^32), every n < 2^31 satisfies
//   floor(n / d) = floor(n m / 2^(31+l)) = umulhi(n, m) >> (l - 1):
// n m / 2^(31+l) = n/d + n (m - 2^(31+l)/d) / 2^(31+l), and the second
// term is below 2^31 / 2^(31+l) = 2^-l < 1/d, less than the distance from
// n/d to the next integer (Granlund and Montgomery, "Division by invariant
// integers using multiplication", 1994, the round-up multiplier).
// A power of two d = 2^k is stored as m = 0, shift k.
#pragma once

#include <stdint.h>

#ifdef __HIPCC__
#define TLBM_HD __host__ __device__
#else
#define TLBM_HD  // plain C++ (host tests)
#endif

namespace tlbm {

struct UDiv {
    uint32_t m, s;
};

TLBM_HD inline UDiv udiv_magic(uint32_t d) {
    uint32_t l = 0;
    while ((uint64_t(1) << l) < uint64_t(d)) ++l;
    if ((uint64_t(1) << l) == uint64_t(d)) return UDiv{0u, l};
    const uint64_t m = ((uint64_t(1) << (31 + l)) + d - 1) / d;
    return UDiv{uint32_t(m), l - 1};
}

// umulhi written as a 64-bit product: hipcc selects v_mul_hi_u32 for it.
TLBM_HD inline uint32_t udiv(uint32_t n, UDiv q) {
    return q.m ? uint32_t((uint64_t(n) * q.m) >> 32) >> q.s : (n >> q.s);
}

}  // namespace tlbm
