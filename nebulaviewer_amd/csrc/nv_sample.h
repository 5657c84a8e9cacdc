// nv_sample.h - the trilinear field sampler of the ray casts and the stirring (sample_n, nebula3d.py:195-226).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "nv_common.h"
#include "nv_divide.h"

namespace nv {

// How a sampler turns a world coordinate into cell units: p / dx through the shared double reciprocal (the
// identical correctly rounded quotient, nv_divide.h), or p * (1/dx) when dx is a power of two, which rounds the
// same real number and so gives the same float. A compile-time choice, so only one form is issued.
struct ScaleDiv {
  Recip rdx;
  __device__ __forceinline__ float operator()(float p) const { return div_r(p, rdx); }
};
struct ScalePow2 {
  float inv;
  __device__ __forceinline__ float operator()(float p) const { return p * inv; }
};

// Loads at a cell index: OFF = uint32_t forms 32-bit byte offsets from the field's base (grids up to 2^30
// cells, so that 4 * index < 2^32, and with nx ny < 2^24), uint64_t 64-bit ones.
//
// The cell index (i ny + j) nz + k in 32 bits: on the uint32_t path the row i ny + j is below 2^24, so both
// products are exact 24-bit multiply-adds (full rate; a 32-bit multiply issues at a quarter of it).
template <class OFF>
__device__ __forceinline__ unsigned cell_index(unsigned i, unsigned j, unsigned k, unsigned ny, unsigned nz) {
  if constexpr (sizeof(OFF) == 4) return __umul24(__umul24(i, ny) + j, nz) + k;
  else return (i * ny + j) * nz + k;
}
template <class OFF>
__device__ __forceinline__ float ld1(const float* __restrict__ f, unsigned i) {
  return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(f) + (OFF)i * 4u);
}
// two z-neighbours (i, i+1) in one 8-byte load (4-byte aligned, which global loads allow)
template <class OFF>
__device__ __forceinline__ float2 ld2(const float* __restrict__ f, unsigned i) {
  typedef float f2u __attribute__((ext_vector_type(2), aligned(4)));
  const f2u v = *reinterpret_cast<const f2u*>(reinterpret_cast<const char*>(f) + (OFF)i * 4u);
  return make_float2(v.x, v.y);
}

// trilinear sample at world p, 0 outside [-1/2, n-1/2] in cell units (sample_n, nebula3d.py:195-226), with
// 32-bit cell indices (grids up to 2^32 cells). PAIR (nz >= 2): the corners at ka = clamp(k0) and
// kb = clamp(k0+1) come from one load of (kp, kp+1), kp = clamp(k0, 0, nz-2): k0 = -1 gives ka = kb = 0 (the
// pair's first value), k0 = nz-1 gives ka = kb = nz-1 (its second), so every corner is the value the separate
// clamped loads read.
template <bool PAIR, class OFF, class SC>
__device__ __forceinline__ float sample_r(const float* __restrict__ f, int nx, int ny, int nz, float px, float py,
                                          float pz, const SC& cs) {
  const float gx = cs(px) - 0.5f, gy = cs(py) - 0.5f, gz = cs(pz) - 0.5f;
  if (gx < -0.5f || gy < -0.5f || gz < -0.5f) return 0.0f;
  if (gx > (float)nx - 0.5f || gy > (float)ny - 0.5f || gz > (float)nz - 0.5f) return 0.0f;
  // (float)(int)floor(g) is floor(g) itself (|g| < 2^24 here), so the fractions come from the float floors
  const float flx = floorf(gx), fly = floorf(gy), flz = floorf(gz);
  const int i0 = (int)flx, j0 = (int)fly, k0 = (int)flz;
  const float fx = gx - flx, fy = gy - fly, fz = gz - flz;
  unsigned r00, r10, r01, r11;
  if constexpr (sizeof(OFF) == 4) {
    // rows i ny + j below 2^24 on this path: clamped in float (one v_med3 each) and formed by one exact fma of
    // integer-valued floats, then a 24-bit multiply (integer clamps of i0 + 1 are rewritten by the compiler into
    // a quarter-rate 32-bit multiply)
    const float xm = (float)(nx - 1), ym = (float)(ny - 1), fny = (float)ny;
    const float fia = fminf(fmaxf(flx, 0.0f), xm), fib = fminf(fmaxf(flx + 1.0f, 0.0f), xm);
    const float fja = fminf(fmaxf(fly, 0.0f), ym), fjb = fminf(fmaxf(fly + 1.0f, 0.0f), ym);
    r00 = __umul24((unsigned)fmaf(fia, fny, fja), nz);
    r10 = __umul24((unsigned)fmaf(fib, fny, fja), nz);
    r01 = __umul24((unsigned)fmaf(fia, fny, fjb), nz);
    r11 = __umul24((unsigned)fmaf(fib, fny, fjb), nz);
  } else {
    const unsigned ia = clampi(i0, 0, nx - 1), ib = clampi(i0 + 1, 0, nx - 1);
    const unsigned ja = clampi(j0, 0, ny - 1), jb = clampi(j0 + 1, 0, ny - 1);
    r00 = cell_index<OFF>(ia, ja, 0, ny, nz);
    r10 = cell_index<OFF>(ib, ja, 0, ny, nz);
    r01 = cell_index<OFF>(ia, jb, 0, ny, nz);
    r11 = cell_index<OFF>(ib, jb, 0, ny, nz);
  }
  float c000, c100, c010, c110, c001, c101, c011, c111;
  if constexpr (PAIR) {
    const unsigned kp = clampi(k0, 0, nz - 2);
    const bool a_second = k0 > nz - 2, b_first = k0 < 0;
    const float2 p00 = ld2<OFF>(f, r00 + kp), p10 = ld2<OFF>(f, r10 + kp);
    const float2 p01 = ld2<OFF>(f, r01 + kp), p11 = ld2<OFF>(f, r11 + kp);
    c000 = a_second ? p00.y : p00.x;
    c100 = a_second ? p10.y : p10.x;
    c010 = a_second ? p01.y : p01.x;
    c110 = a_second ? p11.y : p11.x;
    c001 = b_first ? p00.x : p00.y;
    c101 = b_first ? p10.x : p10.y;
    c011 = b_first ? p01.x : p01.y;
    c111 = b_first ? p11.x : p11.y;
  } else {
    const unsigned ka = clampi(k0, 0, nz - 1), kb = clampi(k0 + 1, 0, nz - 1);
    c000 = ld1<OFF>(f, r00 + ka); c100 = ld1<OFF>(f, r10 + ka); c010 = ld1<OFF>(f, r01 + ka);
    c110 = ld1<OFF>(f, r11 + ka); c001 = ld1<OFF>(f, r00 + kb); c101 = ld1<OFF>(f, r10 + kb);
    c011 = ld1<OFF>(f, r01 + kb); c111 = ld1<OFF>(f, r11 + kb);
  }
  const float c00 = c000 + (c100 - c000) * fx;
  const float c10 = c010 + (c110 - c010) * fx;
  const float c01 = c001 + (c101 - c001) * fx;
  const float c11 = c011 + (c111 - c011) * fx;
  const float c0 = c00 + (c10 - c00) * fy;
  const float c1 = c01 + (c11 - c01) * fy;
  return c0 + (c1 - c0) * fz;
}

// The same trilinear sample from the ghost-layered brick array that nv_ray_bricks builds: brick (p, q, r) holds
// (f(i, ja, ka), f(i, ja, kb), f(i, jb, ka), f(i, jb, kb)) with i = clamp(p-1), ja = clamp(q-1), jb = clamp(q),
// ka = clamp(r-1), kb = clamp(r). A sample inside [-1/2, n-1/2] has i0, j0, k0 in [-1, n-1], and its eight corners
// (ia = clamp(i0), ib = clamp(i0+1), likewise j and k: what sample_r reads) are the bricks at (i0+1, j0+1, k0+1)
// and (i0+2, j0+1, k0+1): two 16-byte loads, no index clamps or selects. Every corner is the value sample_r reads
// and the interpolation is the same expression, so the result is identical.
template <class OFF>
__device__ __forceinline__ float4 ld4(const float4* __restrict__ B, unsigned i) {
  return *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(B) + (OFF)i * 16u);
}

// A brick sample in two parts: brick_pos finds the sample's bricks and weights, brick_interp interpolates the eight
// corners.
struct BrickPos {
  unsigned ia;          // brick (i0+1, j0+1, k0+1); the second brick is ia + (ny+1)(nz+1)
  float fx, fy, fz;
  unsigned keep;        // ~0 inside the box, 0 outside (the sample is +0)
};

template <class OFF, class SC>
__device__ __forceinline__ BrickPos brick_pos(int nx, int ny, int nz, float px, float py, float pz, const SC& cs) {
  const float ux = cs(px) - 0.5f, uy = cs(py) - 0.5f, uz = cs(pz) - 0.5f;
  const float hx = (float)nx - 0.5f, hy = (float)ny - 0.5f, hz = (float)nz - 0.5f;
  const bool outside = ux < -0.5f || uy < -0.5f || uz < -0.5f || ux > hx || uy > hy || uz > hz;
  // No branch: a sample outside the box reads the bricks of its clamped position and returns 0, so the loads of
  // consecutive samples can be in flight together. Inside, the clamp leaves the coordinates unchanged.
  const float gx = fminf(fmaxf(ux, -0.5f), hx), gy = fminf(fmaxf(uy, -0.5f), hy), gz = fminf(fmaxf(uz, -0.5f), hz);
  const float flx = floorf(gx), fly = floorf(gy), flz = floorf(gz);
  const int k0 = (int)flz;
  BrickPos s;
  s.fx = gx - flx;
  s.fy = gy - fly;
  s.fz = gz - flz;
  const unsigned qy = ny + 1, qz = nz + 1;
  if constexpr (sizeof(OFF) == 4) {
    // the padded row (i0 + 1) qy + j0 + 1 is below 2^24 on this path: one exact fma of integer-valued floats,
    // then a 24-bit multiply-add (an integer (i0 + 1) qy, rewritten by the compiler as i0 qy + qy, becomes a
    // quarter-rate 32-bit multiply)
    const unsigned row = (unsigned)fmaf(flx, (float)qy, fly + (float)(qy + 1));
    s.ia = __umul24(row, qz) + (unsigned)(k0 + 1);
  } else {
    s.ia = cell_index<OFF>((int)flx + 1, (int)fly + 1, k0 + 1, qy, qz);
  }
  // +0 outside by a bit mask, not a select: the compiler turns a select over loaded values back into a branch
  s.keep = outside ? 0u : ~0u;
  return s;
}

// the trilinear interpolation of sample_n (nebula3d.py:210-226) from the bricks a = (i0+1, ...), b = (i0+2, ...)
__device__ __forceinline__ float brick_interp(const float4& a, const float4& b, const BrickPos& s) {
  const float c000 = a.x, c100 = b.x, c001 = a.y, c101 = b.y;
  const float c010 = a.z, c110 = b.z, c011 = a.w, c111 = b.w;
  const float c00 = c000 + (c100 - c000) * s.fx;
  const float c10 = c010 + (c110 - c010) * s.fx;
  const float c01 = c001 + (c101 - c001) * s.fx;
  const float c11 = c011 + (c111 - c011) * s.fx;
  const float c0 = c00 + (c10 - c00) * s.fy;
  const float c1 = c01 + (c11 - c01) * s.fy;
  return __uint_as_float(__float_as_uint(c0 + (c1 - c0) * s.fz) & s.keep);
}

template <class OFF, class SC>
__device__ __forceinline__ float sample_brick(const float4* __restrict__ B, int nx, int ny, int nz, float px, float py,
                                              float pz, const SC& cs) {
  const BrickPos s = brick_pos<OFF>(nx, ny, nz, px, py, pz, cs);
  return brick_interp(ld4<OFF>(B, s.ia), ld4<OFF>(B, s.ia + (unsigned)(ny + 1) * (unsigned)(nz + 1)), s);
}

}  // namespace nv
