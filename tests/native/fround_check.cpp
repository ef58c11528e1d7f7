// Compares round_half_away (xnyzip_amd/csrc/fround.hpp) with std::round for
// every float bit pattern.  Built by tests/test_fround.py with g++
// (-ffp-contract=off, SSE arithmetic: IEEE single precision, round to
// nearest even).
//
//   fround_check            all 2^32 patterns
//   fround_check first      stop at the first mismatch
//
// Prints the number of mismatches (NaN inputs must give NaN) and the first
// few, and exits non-zero if any.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "../../xnyzip_amd/csrc/fround.hpp"

int main(int argc, char**) {
  const bool first = argc > 1;
  uint64_t bad = 0;
  for (uint64_t u = 0; u <= 0xFFFFFFFFull; ++u) {
    const uint32_t b = (uint32_t)u;
    float x;
    std::memcpy(&x, &b, 4);
    const float want = std::round(x), got = xnyz::round_half_away(x);
    uint32_t wb, gb;
    std::memcpy(&wb, &want, 4);
    std::memcpy(&gb, &got, 4);
    if (wb == gb || (want != want && got != got)) continue;
    if (bad < 8) std::printf("mismatch: x=%a round=%a fast=%a\n", x, want, got);
    ++bad;
    if (first) break;
  }
  std::printf("%s, %llu mismatches\n", first ? "up to the first mismatch" : "4294967296 patterns",
              (unsigned long long)bad);
  return bad ? 1 : 0;
}
