/* Checks the division scheme of nebulaviewer_amd/csrc/nv_divide.h on the CPU: a float reciprocal with up to
 * +-2 ulp error (the hardware v_rcp_f32 is within 1 ulp), one cubic correction r (1 + e + e^2) in double, one
 * double product rounded to float must equal the IEEE float quotient a/b. Random normal operands over a wide
 * exponent range, quotients placed next to float rounding midpoints, numerators whose quotient is subnormal, and
 * quotients exactly on a midpoint of the subnormal grid (a = b M 2^-150, M odd), which the product alone can
 * round the wrong way. Prints the mismatch count; exit 1 if any. */
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
  const double r = (double)ulps(1.0f / b, seed_err);
  const double e = fma(-d, r, 1.0);
  const float q = (float)((double)a * fma(r, fma(e, e, e), r));
  return fpclassify(q) == FP_SUBNORMAL ? a / b : q;   /* ties of the subnormal grid: IEEE division */
}

int main(int argc, char** argv) {
  long n = argc > 1 ? atol(argv[1]) : 20000000L, bad = 0;
  for (long it = 0; it < n; ++it) {
    float a = mkf(-60, 40, 1), b = mkf(-30, 30, 0);
    if (it % 8 == 2) {
      /* subnormal quotient: a tiny (normal or subnormal), b large enough */
      uint32_t bits = (uint32_t)(rnd() & 0x00ffffffu) | (uint32_t)((rnd() & 1u) << 31);
      memcpy(&a, &bits, 4);
      b = mkf(0, 30, 0);
    } else if (it % 8 == 6) {
      /* a / b an exact tie of the subnormal grid: b = B 2^k, a = B M 2^(k-150) with B M < 2^24 a float */
      uint32_t B = (uint32_t)(rnd() % 4095u) + 2u, M = (uint32_t)(rnd() % 2047u) * 2u + 1u;
      int k = (int)(rnd() % 20);
      if ((uint64_t)B * M >= (1u << 24)) { --it; continue; }
      b = ldexpf((float)B, k);
      a = ldexpf((float)(B * M), k - 150);
      if ((double)a != ldexp((double)B * M, k - 150)) { --it; continue; }
      if (rnd() & 1) a = -a;
    } else if (it % 4 == 3) {
      /* the hardest quotients: B M = K 2^25 + t with B odd (24 bits), M odd (25 bits), |t| < 64, so that
         K / B = M / 2^25 - t / (B 2^25) lies t 2^-49 (relative, about) from the float midpoint M / 2^25 */
      uint32_t B = (uint32_t)(rnd() & 0x7fffffu) | 0x800001u;
      uint32_t inv = B;                                   /* B^-1 mod 2^25 by Newton's iteration */
      for (int k = 0; k < 5; ++k) inv = (inv * (2u - B * inv)) & 0x1ffffffu;
      int32_t t = (int32_t)(rnd() % 32) * 2 + 1;
      if (rnd() & 1) t = -t;
      uint32_t M = ((uint32_t)t * inv) & 0x1ffffffu;
      if (M < 0x1000000u) { --it; continue; }
      uint64_t P = (uint64_t)B * M;
      uint64_t K = (uint64_t)((int64_t)P - t) >> 25;
      int ea = (int)(rnd() % 41) - 20, eb = (int)(rnd() % 41) - 20;
      a = ldexpf((float)K, ea);
      b = ldexpf((float)B, eb - 23);
    } else if (it & 1) {
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
