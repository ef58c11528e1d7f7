// Compares div_const (xnyzip_amd/csrc/exact_div.hpp) with IEEE division on
// the CPU.  Built by tests/test_exact_div.py with g++ (-ffp-contract=off;
// std::fma is the correctly rounded fused multiply-add on every target).
//
//   exact_div_check <divisor-bits-hex>...
//
// For each divisor d: every float of the binades [2^k, 2^(k+1)) for
// k in {-1, 13, 23} (magnitudes of shifted coordinates
// and of recovered nodes), both signs; the dividends around the fast-path
// limits 2^-60 and 2^64; zeros, subnormals, infinities and NaN; and 2^22
// random bit patterns.  Prints the number of mismatching quotients and
// exits non-zero if any.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <initializer_list>

#include "../../xnyzip_amd/csrc/exact_div.hpp"

static float bits_f(uint32_t u) {
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}
static uint32_t f_bits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}

static long check(float a, float d, float r) {
  const float x = a / d, y = xnyz::div_const(a, d, r);
  if (f_bits(x) == f_bits(y) || (x != x && y != y)) return 0;
  std::printf("mismatch: a=%a d=%a ieee=%a fast=%a\n", a, d, x, y);
  return 1;
}

int main(int argc, char** argv) {
  long bad = 0, n = 0;
  for (int i = 1; i < argc; ++i) {
    const float d = bits_f((uint32_t)std::strtoul(argv[i], nullptr, 16));
    if (!xnyz::div_const_ok(d)) {
      std::printf("divisor %a is outside the fast range\n", d);
      return 2;
    }
    const float r = 1.0f / d;
    for (int k : {-1, 13, 23}) {
      const uint32_t lo = f_bits(bits_f(0x3F800000u) * (k >= 0 ? (float)(1u << k) : 0.5f));
      for (uint32_t u = lo; u < lo + (1u << 23); ++u) {
        bad += check(bits_f(u), d, r) + check(-bits_f(u), d, r);
        n += 2;
      }
    }
    for (float e : {0x1p-60f, 0x1p64f}) {
      const uint32_t c = f_bits(e);
      for (uint32_t u = c - 4096; u <= c + 4096; ++u) {
        bad += check(bits_f(u), d, r);
        ++n;
      }
    }
    for (uint32_t u : {0x00000000u, 0x80000000u, 0x00000001u, 0x007FFFFFu, 0x00800000u,
                       0x7F7FFFFFu, 0x7F800000u, 0xFF800000u, 0x7FC00000u}) {
      bad += check(bits_f(u), d, r);
      ++n;
    }
    uint64_t s = 0x9E3779B97F4A7C15ull * (uint64_t)i;
    for (int j = 0; j < (1 << 22); ++j) {
      s ^= s << 13;
      s ^= s >> 7;
      s ^= s << 17;
      bad += check(bits_f((uint32_t)s), d, r);
      ++n;
    }
  }
  std::printf("checked %ld quotients, %ld mismatches\n", n, bad);
  return bad ? 1 : 0;
}
