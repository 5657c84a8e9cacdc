// nv_math.h - deterministic float transcendentals shared by the HIP kernels and the CPU oracle.
//
// The reference kernels call wp.sin / wp.cos (cast_rays nebula3d.py:255-257, cast_fuv nebula3d.py:328-330),
// wp.atan2 (apply_front nebula3d.py:287, apply_fuv nebula3d.py:361) and wp.exp (cast_fuv nebula3d.py:345,351).
// Device math libraries and host libm round these differently in the last place, which would make the GPU
// path and the CPU oracle disagree bit-wise. Both sides therefore use the functions below: float in, float
// out, evaluated with only IEEE +, -, *, / (no FMA: every translation unit that includes this is built with
// -ffp-contract=off), so hipcc and gcc produce identical bits. sin, cos and exp are evaluated in double, which
// makes each result the correctly rounded float except within ~1e-16 relative of a rounding midpoint; atan2 is
// evaluated in float (within 2 ulp).
//
// Methods (textbook, no library code reproduced):
//   sin / cos : Cody-Waite reduction by pi/2 (two-part constant), Taylor series to degree 17/18 on |r|<=pi/4.
//   atan2     : float only (cheaper per cell; within 2 ulp): octant folding, atan(t) = pi/6 +
//               atan((sqrt3 t - 1)/(t + sqrt3)) for t > tan(pi/12) folded into one division, Taylor series of
//               atan to degree 11 on |t| <= tan(pi/12).
//   exp       : Cody-Waite reduction by ln2, Taylor series to degree 13 on |r| <= ln2/2, exact 2^k scaling.
// Valid domain: finite arguments; sincos is used for |x| < 2^20 (ray angles lie in [0, 2pi)).
#pragma once
#include <math.h>

#if defined(__HIPCC__)
#define NV_HD __host__ __device__
#else
#define NV_HD
#endif

namespace nvm {

static constexpr double PIO2_HI = 0x1.921fb544p+0;            // pi/2 truncated to 33 bits (k*PIO2_HI exact)
static constexpr double PIO2_LO = 0x1.0b4611a626331p-34;      // pi/2 - PIO2_HI
static constexpr double TWO_OVER_PI = 0x1.45f306dc9c883p-1;
static constexpr double LN2_HI = 0x1.62e42feep-1;             // ln2 truncated to 32 bits
static constexpr double LN2_LO = 0x1.a39ef35793c76p-33;
static constexpr double INV_LN2 = 0x1.71547652b82fep+0;
static constexpr float PI_F32 = 3.14159265358979323846f;
static constexpr float PIO2_F32 = 1.57079632679489661923f;
static constexpr float PIO6_F32 = 0.52359877559829887308f;
static constexpr float SQRT3_F32 = 1.73205080756887729353f;
static constexpr float TAN_PI_12_F32 = 0.26794919243112270647f;

// sin(r), cos(r) for |r| <= pi/4 (Taylor, Horner in r^2)
NV_HD inline double sin_poly(double r) {
  const double r2 = r * r;
  // p = 1/3! - r^2/5! + r^4/7! - ... - r^14/17!, so that sin r = r - r^3 p
  double p = -1.0 / 355687428096000.0;          // -1/17!
  p = p * r2 + 1.0 / 1307674368000.0;           //  1/15!
  p = p * r2 - 1.0 / 6227020800.0;              // -1/13!
  p = p * r2 + 1.0 / 39916800.0;                //  1/11!
  p = p * r2 - 1.0 / 362880.0;                  // -1/9!
  p = p * r2 + 1.0 / 5040.0;                    //  1/7!
  p = p * r2 - 1.0 / 120.0;                     // -1/5!
  p = p * r2 + 1.0 / 6.0;                       //  1/3!
  return r - (r * r2) * p;
}

NV_HD inline double cos_poly(double r) {
  const double r2 = r * r;
  double p = 1.0 / 6402373705728000.0;          //  1/18!
  p = p * r2 - 1.0 / 20922789888000.0;          // -1/16!
  p = p * r2 + 1.0 / 87178291200.0;             //  1/14!
  p = p * r2 - 1.0 / 479001600.0;               // -1/12!
  p = p * r2 + 1.0 / 3628800.0;                 //  1/10!
  p = p * r2 - 1.0 / 40320.0;                   // -1/8!
  p = p * r2 + 1.0 / 720.0;                     //  1/6!
  p = p * r2 - 1.0 / 24.0;                      // -1/4!
  p = p * r2 + 1.0 / 2.0;                       //  1/2!
  // cos = 1 - r^2 (1/2 - r^2/4! + r^4/6! - ...)
  return 1.0 - r2 * p;
}

NV_HD inline void sincosf_det(float xf, float* s, float* c) {
  const double x = (double)xf;
  const double kd = rint(x * TWO_OVER_PI);
  const double r = (x - kd * PIO2_HI) - kd * PIO2_LO;
  const double sp = sin_poly(r);
  const double cp = cos_poly(r);
  const int q = ((int)kd) & 3;
  double so, co;
  if (q == 0) { so = sp; co = cp; }
  else if (q == 1) { so = cp; co = -sp; }
  else if (q == 2) { so = -sp; co = -cp; }
  else { so = -cp; co = sp; }
  *s = (float)so;
  *c = (float)co;
}

NV_HD inline float sinf_det(float x) { float s, c; sincosf_det(x, &s, &c); return s; }
NV_HD inline float cosf_det(float x) { float s, c; sincosf_det(x, &s, &c); return c; }

// atan2 in float arithmetic only (the per-cell ionize and FUV passes evaluate it once per cell and star, where
// the double form cost a third of the pass). The octant is folded first: t = lo/hi with lo = min(|y|, |x|),
// hi = max(|y|, |x|), so t lies in [0, 1]; for t > tan(pi/12) the identity atan(t) = pi/6 + atan((sqrt3 t - 1) /
// (t + sqrt3)) is applied as (sqrt3 lo - hi) / (lo + sqrt3 hi), one division either way, so the series argument
// satisfies |s| <= tan(pi/12) = 0.268. atan(s) = s + s q with q = -s^2/3 + s^4/5 - ... - s^10/11 (Taylor to
// degree 11: truncation below s^12/13 < 1.1e-8 relative); the correction s q is under 2.4% of s, so its own
// rounding is damped. Then pi/2 - a above the diagonal, pi - a for x < 0, and the sign of y. At most 2 ulp from
// the correctly rounded atan2 (tests/test_oracle_math.py sweeps it against libm); C99 signed-zero conventions.
// atan2f_core takes the one division as div(num, den) (den > 0): IEEE division in atan2f_det, a hardware
// reciprocal in the GPU's fast ionize pass.
template <class Div>
NV_HD inline float atan2f_core(float yf, float xf, Div div) {
  const float ay = fabsf(yf), ax = fabsf(xf);
  const bool steep = ay > ax;
  const float lo = steep ? ax : ay, hi = steep ? ay : ax;
  const bool xneg = signbit(xf) != 0;
  float a;
  if (hi == 0.0f) {
    a = xneg ? PI_F32 : 0.0f;
  } else {
    const bool big = lo > TAN_PI_12_F32 * hi;
    const float num = big ? SQRT3_F32 * lo - hi : lo;
    const float den = big ? lo + SQRT3_F32 * hi : hi;
    const float s = div(num, den);
    const float s2 = s * s;
    float q = -1.0f / 11.0f;
    q = q * s2 + 1.0f / 9.0f;
    q = q * s2 - 1.0f / 7.0f;
    q = q * s2 + 1.0f / 5.0f;
    q = q * s2 - 1.0f / 3.0f;
    q = q * s2;
    float t = s + s * q;
    if (big) t = PIO6_F32 + t;
    a = steep ? PIO2_F32 - t : t;
    if (xneg) a = PI_F32 - a;
  }
  return signbit(yf) ? -a : a;
}

struct IeeeDiv {
  NV_HD float operator()(float a, float b) const { return a / b; }
};
NV_HD inline float atan2f_det(float yf, float xf) { return atan2f_core(yf, xf, IeeeDiv{}); }

// exp(x) for finite x; results below the float range flush to +0 (x < -104), above it to +inf (x > 89).
NV_HD inline float expf_det(float xf) {
  const double x = (double)xf;
  if (x < -104.0) return 0.0f;
  if (x > 89.0) return (float)HUGE_VAL;
  const double kd = rint(x * INV_LN2);
  const double r = (x - kd * LN2_HI) - kd * LN2_LO;
  double p = 1.0 / 6227020800.0;                // 1/13!
  p = p * r + 1.0 / 479001600.0;
  p = p * r + 1.0 / 39916800.0;
  p = p * r + 1.0 / 3628800.0;
  p = p * r + 1.0 / 362880.0;
  p = p * r + 1.0 / 40320.0;
  p = p * r + 1.0 / 5040.0;
  p = p * r + 1.0 / 720.0;
  p = p * r + 1.0 / 120.0;
  p = p * r + 1.0 / 24.0;
  p = p * r + 1.0 / 6.0;
  p = p * r + 1.0 / 2.0;
  p = p * r + 1.0;
  p = p * r + 1.0;
  return (float)ldexp(p, (int)kd);
}

}  // namespace nvm
