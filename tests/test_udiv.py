"""Division by launch-time constants (tlbm_amd/csrc/udiv.h), which the
stream-collide kernels use for node coordinates, against integer division:
every divisor up to 4096 with boundary and random numerators, divisors near
powers of two and up to 2^31 - 1, numerators up to 2^31 - 1."""
import os
import subprocess
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROBE = r"""
#include <cstdio>
#include <cstdint>
#include <random>
#include "udiv.h"
using tlbm::udiv; using tlbm::udiv_magic;
static long long bad = 0, checked = 0;
static void one(uint32_t n, uint32_t d, tlbm::UDiv q) {
    ++checked;
    if (udiv(n, q) != n / d) { if (bad < 5) std::printf("n=%u d=%u got %u\n", n, d, udiv(n, q)); ++bad; }
}
int main() {
    std::mt19937 rng(7);
    const uint32_t top = 0x7fffffffu;
    for (uint32_t d = 1; d <= 4096; ++d) {
        const tlbm::UDiv q = udiv_magic(d);
        for (uint32_t k = 0; k < 64; ++k) {            // around multiples of d
            const uint64_t base = uint64_t(rng() % (top / d)) * d;
            for (int e = -2; e <= 2; ++e) {
                const int64_t n = int64_t(base) + e;
                if (n >= 0 && n <= top) one(uint32_t(n), d, q);
            }
        }
        for (uint32_t n = 0; n < 3 * d + 3; ++n) one(n, d, q);
        one(top, d, q); one(top - 1, d, q);
    }
    for (int sh = 1; sh <= 31; ++sh)
        for (int e = -3; e <= 3; ++e) {
            const int64_t d = (int64_t(1) << sh) + e;
            if (d < 1 || d > top) continue;
            const tlbm::UDiv q = udiv_magic(uint32_t(d));
            for (int k = 0; k < 2000; ++k) one(rng() & top, uint32_t(d), q);
            for (int64_t m = 1; m * d <= top && m < 4096; m = m * 3 + 1) {
                one(uint32_t(m * d), uint32_t(d), q); one(uint32_t(m * d - 1), uint32_t(d), q);
            }
            one(top, uint32_t(d), q);
        }
    for (int k = 0; k < 200000; ++k) {
        const uint32_t d = 1 + (rng() % top), n = rng() & top;
        one(n, d, udiv_magic(d));
    }
    std::printf("%lld %lld\n", checked, bad);
    return 0;
}
"""


def test_udiv_matches_integer_division():
    with tempfile.TemporaryDirectory() as tmp:
        src = os.path.join(tmp, "probe.cpp")
        with open(src, "w") as fh:
            fh.write(PROBE)
        exe = os.path.join(tmp, "probe")
        subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "tlbm_amd", "csrc"),
                        src, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()
    checked, bad = int(out[-2]), int(out[-1])
    assert checked > 1_000_000 and bad == 0, out
