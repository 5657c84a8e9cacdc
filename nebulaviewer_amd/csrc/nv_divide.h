// nv_divide.h - IEEE-exact float division by a shared divisor, two ways: through a double-precision reciprocal
// (Recip, div_r, div_rn: any normal divisor, the ray and render kernels) and float-only (RecipF, div_fn: the
// instruction-bound hydro stage, below).
//
// Double reciprocal, normal quotients. For floats a, b (b normal, positive here) whose quotient x = a/b is a normal
// float or rounds to one, x is never within 2^-49 (relative) of a float rounding midpoint unless it is one: with a =
// A 2^p, b = B 2^q (24-bit A, B) and a midpoint m = M 2^s (M odd, 25 bits), a - b m is a nonzero multiple of 2^min(p,
// q+s) (the odd part of B M exceeds 2^24 > A), which bounds |x - m| / x below by min(2^-24, 2^-49). A double r with
// |r b - 1| <= 2^-52 gives q = RN_double(a r) with relative error < 2^-51.9, so no midpoint separates q from x and
// (float)q == RN_float(a / b): exactly the IEEE division of the reference and of the CPU oracle, but the reciprocal
// is shared by all numerators.
//
// Double reciprocal, subnormal quotients. Below 2^-126 the midpoints are odd multiples of 2^-150 with fewer than 25
// bits, and x can be one exactly (3 2^-149 / 6 = 2^-150, whose IEEE quotient is the even neighbour 0): q, a hair off
// the tie, may round to the other neighbour. Every other x stays at least 2^-25 (relative) from the midpoints (the
// same divisibility argument), so only ties are at stake, and a tie resolved the wrong way always lands on a
// subnormal float (a tie at 2^-150 that rounds down gives 0, which is the IEEE answer). div_r therefore redoes the
// division in IEEE arithmetic when the product rounds to a subnormal: the rare branch, since that needs |a| < 2^-126
// b.
//
// The reciprocal comes from the hardware float reciprocal r0 (<= 1 ulp, 2^-23) and one cubic correction in
// double: e = 1 - d r0 is exact (d r0 has at most 48 significant bits and |e| <= 2^-22), 1/d = r0 (1 + e + e^2 +
// ...), so r = r0 (1 + (e + e^2)) is off by e^3 (2^-66) plus the roundings of e + e^2 (2^-53 |e|) and of the last
// fma (2^-53): |r d - 1| < 2^-52. Valid for normal divisors (the argument is symmetric in sign), which is
// what the kernels divide by (density floor > 0, dx > 0, SR - SL > 0, constants).
#pragma once
#include <hip/hip_runtime.h>

namespace nv {

// a float divisor b with its shared reciprocal r (recip_d(b), or the correctly rounded 1.0 / (double)b)
struct Recip {
  double r;
  float b;
};

__device__ __forceinline__ double recip_d(float b) {
  const double d = (double)b;
  const double r = (double)__builtin_amdgcn_rcpf(b);
  const double e = fma(-d, r, 1.0);
  return fma(r, fma(e, e, e), r);
}

__device__ __forceinline__ Recip recip(float b) { return Recip{recip_d(b), b}; }

constexpr int SUBNORMAL = 0x0090;   // __builtin_isfpclass mask: negative or positive subnormal

// RN_float(a / R.b)
__device__ __forceinline__ float div_r(float a, const Recip& R) {
  const float q = (float)((double)a * R.r);
  if (__builtin_expect(__builtin_isfpclass(q, SUBNORMAL), 0)) return a / R.b;
  return q;
}

// q[i] = RN_float(a[i] / R.b) for N numerators of one divisor, with one branch for the group: if any product
// rounds to a subnormal, all N are redone in IEEE arithmetic
template <int N>
__device__ __forceinline__ void div_rn(const float (&a)[N], const Recip& R, float (&q)[N]) {
  bool sub = false;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    q[i] = (float)((double)a[i] * R.r);
    sub = sub || __builtin_isfpclass(q[i], SUBNORMAL);
  }
  if (__builtin_expect(sub, 0)) {
#pragma unroll
    for (int i = 0; i < N; ++i) q[i] = a[i] / R.b;
  }
}

// ------------------------------------------------------------------------------------------------------------
// Float-only exact division (Markstein's scheme), for the instruction-bound stage.
//
// y = RN(1/b) from the hardware reciprocal y0 and one Newton step fma(fma(-b, y0, 1), y0, y0); q0 = RN(a y);
// r = b q0 - a; q = RN(q0 - r y). Two exhaustive runs on gfx950 establish it: y equals RN(1/b) for every
// normal b below 2^126, and q equals the IEEE quotient for all 2^46 pairs of significands a, b in [1, 2). The
// sequence is a scaled copy of that case whenever nothing leaves the normal range, which the guard below
// ensures: b in [2^-37, 2^91) (y normal) and |q0| in [2^-64, 2^64) (q0 and q normal, no overflow, and
// |a| >= 2^-102). The residual r is a multiple of 2^(e_a - 47) (the lowest bits of a and of b q0), so one that
// fits 24 bits is a multiple of 2^-149, exact at any scale, and one that must be rounded has at least 25 bits,
// |r| >= 2^(e_a - 23) >= 2^-125, and rounds as in [1, 2). The divisor's range is tested with the quotients:
// t = b 2^-27 lies in [2^-64, 2^64) exactly when b lies in [2^-37, 2^91), so one min/max chain over |q0| and t
// covers both (an infinite b makes t infinite; a NaN b gives NaN quotients either way). Outside the guard the
// group is redone in IEEE arithmetic; zero numerators stay on the fast path (r = +0 and q keeps the sign of a/b).
struct RecipF {
  float y, b;
  float t;   // b 2^-27
};

__device__ __forceinline__ RecipF recip_f(float b) {
  const float y0 = __builtin_amdgcn_rcpf(b);
  return RecipF{fmaf(fmaf(-b, y0, 1.0f), y0, y0), b, b * 0x1p-27f};
}

// q[i] = RN_float(a[i] / R.b) for N numerators of one divisor; one guard for the group
template <int N>
__device__ __forceinline__ void div_fn(const float (&a)[N], const RecipF& R, float (&q)[N]) {
  float q0[N];
  float lo = fabsf(R.t), hi = fabsf(R.t);
#pragma unroll
  for (int i = 0; i < N; ++i) {
    q0[i] = a[i] * R.y;
    const float r = fmaf(R.b, q0[i], -a[i]);
    q[i] = fmaf(-r, R.y, q0[i]);
    lo = fminf(lo, fabsf(q0[i]));
    hi = fmaxf(hi, fabsf(q0[i]));
  }
  if (__builtin_expect(!(lo >= 0x1p-64f && hi < 0x1p64f), 0)) {
    // zeros (common: a fluid at rest) keep the fast result; anything else outside the guard takes IEEE
    const float tb = fabsf(R.t);
    const bool bad_b = !(tb >= 0x1p-64f && tb < 0x1p64f);
    bool need[N], any = false;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const float m = fabsf(q0[i]);
      need[i] = bad_b || (a[i] != 0.0f && !(m >= 0x1p-64f && m < 0x1p64f));
      any = any || need[i];
    }
    if (any) {
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const float x = a[i] / R.b;
        q[i] = need[i] ? x : q[i];
      }
    }
  }
}

// Correctly rounded square root: the hardware square root s (<= 1 ulp) moved by one ulp where the residuals of
// its neighbours say so (the IEEE expansion of sqrtf without its rescaling of small arguments). Exhaustively
// verified against the IEEE square root on gfx950 for every float from 2^-96 up; smaller arguments take sqrtf.
__device__ __forceinline__ float sqrt_rn(float x) {
  if (__builtin_expect(x < 0x1p-96f, 0)) return sqrtf(x);
  const float s = __builtin_amdgcn_sqrtf(x);
  const float sm = __uint_as_float(__float_as_uint(s) - 1u), sp = __uint_as_float(__float_as_uint(s) + 1u);
  const float rm = fmaf(-sm, s, x), rp = fmaf(-sp, s, x);
  const float t = rm <= 0.0f ? sm : s;
  return rp > 0.0f ? sp : t;
}

}  // namespace nv
