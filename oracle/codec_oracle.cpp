// codec_oracle.cpp - CPU restatement of the kernels around the step that consumers run on its fields: the
// timeline snapshot codec, the shock proxy, and the render volume packing. TEST INFRASTRUCTURE ONLY (loaded by
// tests/ as the checker; the product never links it). Float arithmetic in the reference's operation order,
// built with -ffp-contract=off like the rest of the oracle.
//
// Pinning (tests/test_codec_oracle.py): the codec's round trip is exact on values binary16 / 8 bits represent
// and within half a quantum otherwise, saturates at +-65000 and at the density floor; the shock proxy is zero for
// uniform and expanding flows and equals -div(v) rho on a linear compression (central differences are exact on
// linear fields); the pack equals hand-evaluated emission on single cells (sstep edge values, the fade profile,
// the G0 saturation); smooth_axis maps constants to themselves and a delta to [1 4 1]/6 (then binary16).
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <algorithm>

namespace {

inline int iclamp(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// IEEE binary32 -> binary16, round to nearest even (wp.float16(x))
inline uint16_t to_h(float f) {
  uint32_t x;
  memcpy(&x, &f, 4);
  uint32_t sign = (x >> 16) & 0x8000u;
  uint32_t ax = x & 0x7fffffffu;
  if (ax >= 0x7f800000u) return (uint16_t)(sign | 0x7c00u | (ax > 0x7f800000u ? 0x200u : 0u));
  if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);
  if (ax < 0x38800000u) {
    if (ax < 0x33000000u) return (uint16_t)sign;
    uint32_t e = ax >> 23;
    uint32_t m = (ax & 0x7fffffu) | 0x800000u;
    uint32_t shift = 126u - e;
    uint32_t q = m >> shift;
    uint32_t rem = m & ((1u << shift) - 1u);
    uint32_t half = 1u << (shift - 1u);
    if (rem > half || (rem == half && (q & 1u))) ++q;
    return (uint16_t)(sign | q);
  }
  uint32_t e = (ax >> 23) - 112u;
  uint32_t m = ax & 0x7fffffu;
  uint32_t q = (e << 10) | (m >> 13);
  uint32_t rem = m & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (q & 1u))) ++q;
  return (uint16_t)(sign | q);
}

// binary16 -> binary32 (exact)
inline float from_h(uint16_t h) {
  uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
  uint32_t e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
  uint32_t x;
  if (e == 0x1fu) {
    x = sign | 0x7f800000u | (m << 13);
  } else if (e == 0) {
    if (m == 0) {
      x = sign;
    } else {                       // subnormal: normalise
      int s = 0;
      while (!(m & 0x400u)) {
        m <<= 1;
        ++s;
      }
      m &= 0x3ffu;
      x = sign | ((uint32_t)(113 - s) << 23) | (m << 13);
    }
  } else {
    x = sign | ((e + 112u) << 23) | (m << 13);
  }
  float f;
  memcpy(&f, &x, 4);
  return f;
}

inline float fminf_(float a, float b) { return a < b ? a : b; }
inline float fmaxf_(float a, float b) { return a > b ? a : b; }
inline float clampf_(float x, float lo, float hi) { return fminf_(fmaxf_(x, lo), hi); }

struct G3 {
  int nx, ny, nz;
  size_t at(int i, int j, int k) const { return ((size_t)i * ny + j) * nz + k; }
  size_t atc(int i, int j, int k) const {
    return at(iclamp(i, 0, nx - 1), iclamp(j, 0, ny - 1), iclamp(k, 0, nz - 1));
  }
};

// shock proxy at one cell (nebula3d.py:182-191, render.py:166-174): max(0, -div v) rho, central differences of
// v = m / rho with clamped neighbours
inline float shock_at(const float* rho, const float* mx, const float* my, const float* mz, const G3& g, int i, int j,
                      int k, float dx) {
  const float du = mx[g.atc(i + 1, j, k)] / rho[g.atc(i + 1, j, k)] - mx[g.atc(i - 1, j, k)] / rho[g.atc(i - 1, j, k)];
  const float dv = my[g.atc(i, j + 1, k)] / rho[g.atc(i, j + 1, k)] - my[g.atc(i, j - 1, k)] / rho[g.atc(i, j - 1, k)];
  const float dw = mz[g.atc(i, j, k + 1)] / rho[g.atc(i, j, k + 1)] - mz[g.atc(i, j, k - 1)] / rho[g.atc(i, j, k - 1)];
  const float div = (du + dv + dw) / (2.0f * dx);
  return fmaxf_(-div, 0.0f) * rho[g.at(i, j, k)];
}

// render.py:121-124, 127-130
inline float sstep(float e0, float e1, float x) {
  const float t = clampf_((x - e0) / (e1 - e0), 0.0f, 1.0f);
  return t * t * (3.0f - 2.0f * t);
}
inline float fade_axis(float u, float e) {
  if (e <= 0.0f) return 1.0f;
  return sstep(0.0f, e, u) * sstep(0.0f, e, 1.0f - u);
}

}  // namespace

extern "C" {

// timeline.k_pack / k_pack_lm (timeline.py:19-34, 57-72): dfront and g0 given as float values (the binary16
// fields of k_pack_lm are converted exactly first, which is what float(df[i, j, k]) does)
void or_timeline_pack(const float* rho, const float* mx, const float* my, const float* mz, const float* xi,
                      const float* df, const float* g0, uint16_t* o_r, uint16_t* o_mx, uint16_t* o_my, uint16_t* o_mz,
                      uint8_t* o_xi, uint16_t* o_df, uint16_t* o_g0, long long n) {
  const float m = 65000.0f;
  for (long long q = 0; q < n; ++q) {
    o_r[q] = to_h(fminf_(rho[q], m));
    o_mx[q] = to_h(clampf_(mx[q], -m, m));
    o_my[q] = to_h(clampf_(my[q], -m, m));
    o_mz[q] = to_h(clampf_(mz[q], -m, m));
    o_xi[q] = (uint8_t)(int)(clampf_(xi[q], 0.0f, 1.0f) * 255.0f + 0.5f);
    o_df[q] = to_h(clampf_(df[q], -m, m));
    o_g0[q] = to_h(fminf_(g0[q], m));
  }
}

// timeline.k_unpack / k_unpack_lm (timeline.py:37-52, 75-89): dfront / g0 come back as float values
void or_timeline_unpack(const uint16_t* s_r, const uint16_t* s_mx, const uint16_t* s_my, const uint16_t* s_mz,
                        const uint8_t* s_xi, const uint16_t* s_df, const uint16_t* s_g0, float floor_, float* rho,
                        float* mx, float* my, float* mz, float* xi, float* df, float* g0, long long n) {
  for (long long q = 0; q < n; ++q) {
    rho[q] = fmaxf_(from_h(s_r[q]), floor_);
    mx[q] = from_h(s_mx[q]);
    my[q] = from_h(s_my[q]);
    mz[q] = from_h(s_mz[q]);
    xi[q] = (float)(int)s_xi[q] / 255.0f;
    df[q] = from_h(s_df[q]);
    g0[q] = from_h(s_g0[q]);
  }
}

// nebula3d.shock_field (nebula3d.py:182-191)
void or_shock_field(const float* rho, const float* mx, const float* my, const float* mz, int nx, int ny, int nz,
                    float dx, float* out) {
  const G3 g{nx, ny, nz};
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) out[g.at(i, j, k)] = shock_at(rho, mx, my, mz, g, i, j, k, dx);
}

// render.pack_volume_lm (render.py:176-204): out = binary16 (em.r, em.g, em.b, dust) per cell. pp = c_ha[3],
// c_o3[3], c_s2[3], c_dg[3], c_sh[3], gs, ambient, edge, inv_L[3], dx (the PackParams order, render.py:77-88).
void or_pack_volume(const float* rho, const float* mx, const float* my, const float* mz, const float* xi,
                    const float* df, const float* g0, int nx, int ny, int nz, const float* pp, int shock_on,
                    uint16_t* out) {
  const G3 g{nx, ny, nz};
  const float *c_ha = pp, *c_o3 = pp + 3, *c_s2 = pp + 6, *c_dg = pp + 9, *c_sh = pp + 12;
  const float gs = pp[15], ambient = pp[16], edge = pp[17], dx = pp[21];
  const float* inv_L = pp + 18;
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) {
        const size_t o = g.at(i, j, k);
        const float r = rho[o];
        const float x = clampf_(xi[o], 0.0f, 1.0f);
        const float d = df[o];
        const float ni = r * x;
        const float ha = ni * ni;
        const float oiii = ha * sstep(-0.03f, -0.12f, d);
        const float sii = ha * sstep(-0.05f, -0.005f, d);
        float dust = r * (1.0f - x) + 0.08f * ni;
        const float gg = g0[o];
        const float geff = gg / (1.0f + gg / gs) + ambient;
        const float sh = shock_on != 0 ? shock_at(rho, mx, my, mz, g, i, j, k, dx) : 0.0f;
        const float p[3] = {(((float)i + 0.5f) * dx) * inv_L[0], (((float)j + 0.5f) * dx) * inv_L[1],
                            (((float)k + 0.5f) * dx) * inv_L[2]};
        const float f = fade_axis(p[0], edge) * fade_axis(p[1], edge) * fade_axis(p[2], edge);
        float em[3];
        for (int c = 0; c < 3; ++c) {
          em[c] = ha * c_ha[c] + oiii * c_o3[c] + sii * c_s2[c] + (dust * geff) * c_dg[c] + sh * c_sh[c];
          em[c] = em[c] * f;
        }
        dust = dust * f;
        const float m = 60000.0f;
        out[4 * o + 0] = to_h(fminf_(em[0], m));
        out[4 * o + 1] = to_h(fminf_(em[1], m));
        out[4 * o + 2] = to_h(fminf_(em[2], m));
        out[4 * o + 3] = to_h(fminf_(dust, m));
      }
}

// render.smooth_axis (render.py:281-298): [1 4 1]/6 along one axis of a binary16 RGBA grid, clamped ends
void or_smooth_axis(const uint16_t* src, uint16_t* dst, int axis, int nx, int ny, int nz) {
  const G3 g{nx, ny, nz};
  const int oi = axis == 0, oj = axis == 1, ok = axis == 2;
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) {
        const size_t a = g.at(std::max(i - oi, 0), std::max(j - oj, 0), std::max(k - ok, 0));
        const size_t b = g.at(i, j, k);
        const size_t c = g.at(std::min(i + oi, nx - 1), std::min(j + oj, ny - 1), std::min(k + ok, nz - 1));
        for (int ch = 0; ch < 4; ++ch)
          dst[4 * b + ch] =
              to_h((from_h(src[4 * a + ch]) + 4.0f * from_h(src[4 * b + ch]) + from_h(src[4 * c + ch])) / 6.0f);
      }
}

}  // extern "C"
