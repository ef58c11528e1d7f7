// exact_div.hpp -- correctly rounded division by a constant divisor.
//
// The quantizers divide every coordinate by one divisor per cloud
// (truncated_octahedron_quantizer.hpp:29, cube_quantizer.hpp:20) and the
// recover step by float(sqrt(5)) or float(sqrt(3))
// (truncated_octahedron_quantizer.hpp:72, cube_quantizer.hpp:34); the lattice
// nodes must equal the IEEE quotient bit for bit.  With r = RN(1/d) computed
// once on the host, a / d is formed by a product and two FMA corrections:
//
//   q0 = RN(a r);  q1 = q0 + RN(a - q0 d) r;  q  = q1 + RN(a - q1 d) r
//
// Markstein's theorem (P. Markstein, "Computation of elementary functions on
// the IBM RISC System/6000 processor", IBM J. Res. Dev. 34(1), 1990; also
// Muller et al., Handbook of Floating-Point Arithmetic, the division
// chapter): when r is within half an ulp of 1/d and q1 within one ulp of
// a/d, the residual a - q1 d is exact and q is RN(a/d).  q0 is within two
// ulps and the first correction brings q1 within one.  The theorem assumes
// no underflow or overflow in the steps, so the fast form is used only for
// divisors in [2^-40, 2^40] (checked on the host) and dividends whose
// magnitude is in [2^-60, 2^64]; zeros keep their sign, anything else (tiny,
// huge, infinite or NaN dividends) takes the IEEE division.  tests/test_exact_div.py compares
// it with IEEE division on the CPU (g++, fmaf) for the divisors of many
// bounds, whole binades of dividends and their edges.
#pragma once

#if defined(__HIPCC__)
#define XNYZ_HD __host__ __device__
#else
#define XNYZ_HD
#endif

#include <cmath>

namespace xnyz {

// Whether d's reciprocal may stand in for it (host side, once per cloud).
inline bool div_const_ok(float d) {
  return std::isfinite(d) && d >= 0x1p-40f && d <= 0x1p40f;
}

// RN(a / d) given r = RN(1 / d) and div_const_ok(d).
XNYZ_HD inline float div_const(float a, float d, float r) {
  const float x = std::fabs(a);
  if (!(x >= 0x1p-60f && x <= 0x1p64f)) return a == 0.0f ? a * r : a / d;  // keeps -0
  const float q0 = a * r;
  const float q1 = std::fma(std::fma(-q0, d, a), r, q0);
  return std::fma(std::fma(-q1, d, a), r, q1);
}

}  // namespace xnyz
