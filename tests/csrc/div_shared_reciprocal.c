/* Checks the division scheme of nebulaviewer_amd/csrc/nv_divide.h on the CPU: a float reciprocal with up to
 * +-2 ulp error (the hardware v_rcp_f32 is within 1 ulp), two double-precision fma Newton steps, one double
 * product rounded to float must equal the IEEE float quotient a/b. Random normal operands over a wide exponent
 * range, plus quotients placed next to float rounding midpoints. Prints the mismatch count; exit 1 if any. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static uint64_t s = 88172645463325252ull;
static uint64_t rnd(void) { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return s; }
static float mkf(int emin, int emax, int neg) {
  uint32_t m = (uint32_t)(rnd() & 0x7fffff);
  int e = emin + (int)(rnd() % (uint64_t)(emax - emin + 1));
  uint32_t b = ((uint32_t)(e + 127) << 23) | m | (neg ? (uint32_t)((rnd() & 1u) << 31) : 0u);
  float f;
  memcpy(&f, &b, 4);
  return f;
}
static float ulps(float x, int k) { uint32_t b; memcpy(&b, &x, 4); b += (uint32_t)k; memcpy(&x, &b, 4); return x; }

static float div_shared(float a, float b, int seed_err) {
  const double d = (double)b;
  double r = (double)ulps(1.0f / b, seed_err);
  double e = fma(-d, r, 1.0);
  r = fma(r, e, r);
  e = fma(-d, r, 1.0);
  r = fma(r, e, r);
  return (float)((double)a * r);
}

int main(int argc, char** argv) {
  long n = argc > 1 ? atol(argv[1]) : 20000000L, bad = 0;
  for (long it = 0; it < n; ++it) {
    float a = mkf(-60, 40, 1), b = mkf(-30, 30, 0);
    if (it & 1) {
      /* a = b * m rounded, m a float midpoint: the quotient lies within a few ulps of the midpoint */
      float m0 = mkf(-10, 10, 0);
      double mid = (double)m0 + ldexp(1.0, ilogb(m0) - 24);
      a = (float)(mid * (double)b);
    }
    float q = div_shared(a, b, (int)(rnd() % 5) - 2);
    if (q != a / b) {
      if (bad < 5) printf("mismatch a=%a b=%a got=%a want=%a\n", a, b, q, a / b);
      ++bad;
    }
  }
  printf("checked %ld mismatches %ld\n", n, bad);
  return bad ? 1 : 0;
}
