// fround.hpp -- roundf (round half away from zero) in three operations.
//
// The truncated-octahedron quantizer rounds six values per point with
// std::round (truncated_octahedron_quantizer.hpp:31-36).  The device's roundf
// takes a truncation, a subtraction, a magnitude comparison, a select, a
// sign copy and an addition.  This form adds the largest float below one
// half, with x's sign, and truncates:
//
//   round(x) = trunc(x + copysign(0.5 - 2^-25, x))
//
// For |x| < 2^23 the rounded sum reaches the next integer away from zero
// exactly when the fraction of |x| is at least one half (one half itself:
// 0.5 + 0.5 - 2^-25 ties between 1 - 2^-24 and 1 and goes to the even 1;
// above one, k + 1 - 2^-25 is within a quarter ulp of k + 1).  For |x| >= 2^23
// x is an integer and the added value is under half an ulp, so the sum is x.
// Zeros keep their sign; infinities and NaN pass through.
// tests/test_fround.py compares it with std::round for every one of the 2^32
// float bit patterns.
#pragma once

#if defined(__HIPCC__)
#define XNYZ_HD_FR __host__ __device__
#else
#define XNYZ_HD_FR
#endif

namespace xnyz {

XNYZ_HD_FR inline float round_half_away(float x) {
  return __builtin_truncf(x + __builtin_copysignf(0x1.fffffep-2f, x));
}

}  // namespace xnyz
