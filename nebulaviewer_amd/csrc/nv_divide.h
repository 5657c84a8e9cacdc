// nv_divide.h - IEEE-exact float division by a shared divisor, through a double-precision reciprocal.
//
// For floats a, b (b normal, positive here) the exact quotient x = a/b is never within 2^-49 (relative) of a
// float rounding midpoint: with a = A 2^p, b = B 2^q (24-bit A, B) and a midpoint m = M 2^s (M odd, 25 bits),
// a - b m is a nonzero multiple of 2^min(p, q+s) (the odd part of B M exceeds 2^24 > A), which bounds
// |x - m| / x below by min(2^-24, 2^-49). A double r with |r b - 1| <= 2^-52 gives q = RN_double(a r) with
// relative error < 2^-51.9, so no midpoint separates q from x and (float)q == RN_float(a / b): exactly the
// IEEE division of the reference and of the CPU oracle, but the reciprocal is shared by all numerators.
// r comes from the hardware float reciprocal (<= 1 ulp, 2^-23) and two fma Newton steps in double
// (error 2^-46, then below 2^-53 + 2^-92). Valid for normal positive divisors and quotients within the float
// range, which is what the hydro stage divides by (density floor > 0, dx > 0, SR - SL > 0).
#pragma once
#include <hip/hip_runtime.h>

namespace nv {

__device__ __forceinline__ double recip_d(float b) {
  const double d = (double)b;
  double r = (double)__builtin_amdgcn_rcpf(b);
  double e = fma(-d, r, 1.0);
  r = fma(r, e, r);
  e = fma(-d, r, 1.0);
  return fma(r, e, r);
}

// RN_float(a / b) given r = recip_d(b)
__device__ __forceinline__ float div_r(float a, double r) { return (float)((double)a * r); }

}  // namespace nv
