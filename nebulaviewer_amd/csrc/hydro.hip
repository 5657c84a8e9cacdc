// hydro.hip - the fused finite-volume RK stage and the CFL signal-speed reduction for gfx950.
//
// Reference: hydro_stage_hll (sim.py:423-498), nebula3d.hydro_stage (nebula3d.py:76-151), flux_hll
// (sim.py:397-420), face_flux (nebula3d.py:54-72), limiters (nebula3d.py:40-51, sim.py:366-394),
// max_speed (nebula3d.py:154-162).
//
// Layout: structure-of-arrays fields, C-order [nx][ny][nz] (z fastest), as the reference stores them.
//
// Work decomposition. A workgroup owns a TK x TJ column of cells in (z, y) - one wavefront per y row, the 64
// lanes along the contiguous z axis - and marches along x over `ci` planes. Per plane:
//   * every thread keeps its own cell column's primitives, c^2, c and conserved state for the planes it needs
//     in registers (a sliding window: the x stencil needs no neighbours and no halo), so the divisions of
//     nebula3d.prim and the square root of c^2 happen once per cell;
//   * the plane's primitives go to LDS with a 2-cell y/z halo (the halo cells are the only ones converted
//     twice: (TJ+4)(TK+4) - TJ TK of them per TJ TK cells);
//   * every cell computes its MUSCL edge states once per axis (the limited slope is shared by its two faces);
//   * each face flux is evaluated once: x in registers (the right face of plane i is the left face of plane
//     i+1), the low y and z faces of every cell into LDS, plus the TK + TJ faces on the tile's high boundary.
// The reference evaluates six face fluxes per cell (each face twice) from 63 clamped loads; the arithmetic per
// face is the same.
//
// Two compile-time policies:
//   Arithmetic  Exact: no FMA contraction (-ffp-contract=off), IEEE square root, and every division IEEE or
//               through the shared double reciprocal of nv_divide.h (the identical correctly rounded
//               quotient): results equal the CPU oracle bit for bit (up to the sign of zero, which never feeds
//               a nonzero result here).
//               Fast: the same expressions with explicit fused multiply-adds, the hardware reciprocal and
//               square root (each within about one ulp), the median form of minmod, a branchless HLL and a
//               reciprocal-square-root wind term.
//   Storage     F32: (rho, mx, my, mz) float32, the reference layout.
//               F16V: rho float32, (vx, vy, vz) binary16. |v| <= vmax lies in the normal fp16 range, so each
//               velocity keeps a relative precision of 2^-11 with no scale factor; the primitive state is read
//               directly (no division) and the conserved state is rho * v. Density stays float32: rounding it
//               to binary16 every stage stalls updates below half an fp16 ulp (measured: rel-L1 density error
//               4e-3 after 100 steps of the 64^3 torus against 1.4e-4 with float32 density).
#include <string.h>
#include <map>
#include <mutex>
#include <utility>
#include "nv_common.h"
#include "nv_divide.h"

namespace nv {

// Tile shapes (z extent TK, y rows TJ, one thread per cell column): 64 x 8 (one wavefront per y row), 32 x 16
// (one wavefront per two rows), the latter for z extents that are multiples of 32 but not of 64 and where its
// smaller halo wins, and 64 x 16 (1024 threads: half the halo share of 64 x 8, one workgroup per CU).
constexpr int TILE_AUTO = 0, TILE_64x8 = 1, TILE_32x16 = 2, TILE_64x16 = 3;
template <int TK_, int TJ_>
struct Tile {
  static constexpr int TK = TK_, TJ = TJ_, NT = TK * TJ;
  static constexpr int PK = TK + 4, PJ = TJ + 4;     // primitive plane with a 2-cell halo
  static constexpr int CK = TK + 2, CJ = TJ + 2;     // c^2 / c plane with a 1-cell halo
  static constexpr int NHALO = 4 * PK + 4 * TJ;      // halo cells of the primitive plane
  static constexpr int NCHALO = 2 * CK + 2 * TJ;     // halo cells of the c plane
};
using T64 = Tile<64, 8>;
using T32 = Tile<32, 16>;
using T6416 = Tile<64, 16>;

struct HydroArgs {
  const void *f0, *f1, *f2, *f3;            // U: rho and momentum (F32) or rho and velocity (F16V)
  const float* xi;
  const void *g0, *g1, *g2, *g3;            // U0, same storage
  void *o0, *o1, *o2, *o3;                  // output, same storage
  const float4 *stars, *jets;
  float* smax;
  int nx, ny, nz, nstars, ci, u0_is_u;
  int tiles_k, tiles_j, tiles, per_xcd;     // tile grid and the number of tiles each XCD takes
  int run;                                  // tiles an XCD takes in a row before the next XCD's (band_tile)
  float wind, jet_cos, jet_r, pulse, dt, dx, a, b, floor_, cn2, ci2, vmax;
  float vmax2_lo;                           // largest float <= vmax^2: |v|^2 <= vmax2_lo means no cap
  float inv_dx;                             // 1/dx when dx is a power of two (exact), else 0
  float jet_r2_hi;                          // a float above jet_r^2 (real): |r|^2 beyond it means r >= jet_r
  const float* dt_dev;                      // if non-null, dt is read here when the stage runs
};

// ------------------------------------------------------------------------------------------ arithmetic
struct Exact {
  static constexpr bool FAST = false;
  using Rcp = RecipF;
  static __device__ __forceinline__ Rcp rcp(float b) { return recip_f(b); }
  static __device__ __forceinline__ float div(float a, const Rcp& r) {
    const float x[1] = {a};
    float q[1];
    div_fn<1>(x, r, q);
    return q[0];
  }
  template <int N>
  static __device__ __forceinline__ void divn(const float (&a)[N], const Rcp& r, float (&q)[N]) { div_fn<N>(a, r, q); }
  static __device__ __forceinline__ float mad(float a, float b, float c) { return a * b + c; }
  static __device__ __forceinline__ float sqrt(float x) { return sqrt_rn(x); }
};

struct Fast {
  static constexpr bool FAST = true;
  using Rcp = float;
  static __device__ __forceinline__ Rcp rcp(float b) { return __builtin_amdgcn_rcpf(b); }   // <= 1 ulp
  static __device__ __forceinline__ float div(float a, Rcp r) { return a * r; }
  template <int N>
  static __device__ __forceinline__ void divn(const float (&a)[N], Rcp r, float (&q)[N]) {
#pragma unroll
    for (int i = 0; i < N; ++i) q[i] = a[i] * r;
  }
  static __device__ __forceinline__ float mad(float a, float b, float c) { return fmaf(a, b, c); }
  static __device__ __forceinline__ float sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
};

// ------------------------------------------------------------------------------------------ storage
__device__ __forceinline__ float h2f(uint16_t h) { return (float)__builtin_bit_cast(_Float16, h); }

// Addressing. Field element (plane po, in-plane cell c) is at f + po + c. Each access goes through a buffer
// resource whose base is the plane's first element: po is the same for the whole workgroup, so the base is scalar
// arithmetic, and the cell's byte offset within the plane, cb = 4 c (a 2-byte field's is cb / 2), is the one
// 32-bit vector operand (ny nz < 2^30). No 64-bit vector address arithmetic per access, and the per-thread offset
// is one register for every field. The resource's range is the plane (kPlaneBytes4: 4 ny nz bytes of a 4-byte
// field), which every in-plane offset lies inside.
constexpr int kBufWord3 = 0x00020000;   // gfx9 buffer resource word 3: 32-bit data, no swizzle
template <class T>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const void* f, long long po, unsigned bytes4) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(static_cast<const T*>(f) + po), 0,
                                           (int)(sizeof(T) == 4 ? bytes4 : bytes4 >> 1), kBufWord3);
}
// aux: cache policy of the access, 0 (default) or 2 (non-temporal)
template <class T, int AUX = 0>
__device__ __forceinline__ T ldp(const void* f, long long po, unsigned cb, unsigned bytes4) {
  const __amdgpu_buffer_rsrc_t r = plane_rsrc<T>(f, po, bytes4);
  if constexpr (sizeof(T) == 4) return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, cb, 0, AUX));
  else return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b16(r, cb >> 1, 0, AUX));
}

// The stage's outputs and its U0 are touched by one cell only, so they are stored and loaded non-temporally: they
// do not take L2 lines from the state and xi planes, which neighbouring tiles read again as their halos.
constexpr int kNT = 2;
template <class T>
__device__ __forceinline__ void st_once(void* f, long long po, unsigned cb, unsigned bytes4, T v) {
  const __amdgpu_buffer_rsrc_t r = plane_rsrc<T>(f, po, bytes4);
  if constexpr (sizeof(T) == 4) __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, cb, 0, kNT);
  else __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(uint16_t, v), r, cb >> 1, 0, kNT);
}
template <class T>
__device__ __forceinline__ T ld_once(const void* f, long long po, unsigned cb, unsigned bytes4) {
  return ldp<T, kNT>(f, po, cb, bytes4);
}

struct F32 {
  static constexpr bool VEL = false;   // stores momentum
  static __device__ __forceinline__ float4 load(const void* f0, const void* f1, const void* f2, const void* f3,
                                                long long po, unsigned cb, unsigned b4) {
    return make_float4(ldp<float>(f0, po, cb, b4), ldp<float>(f1, po, cb, b4), ldp<float>(f2, po, cb, b4),
                       ldp<float>(f3, po, cb, b4));
  }
  static __device__ __forceinline__ float4 load_once(const void* f0, const void* f1, const void* f2, const void* f3,
                                                     long long po, unsigned cb, unsigned b4) {
    return make_float4(ld_once<float>(f0, po, cb, b4), ld_once<float>(f1, po, cb, b4), ld_once<float>(f2, po, cb, b4),
                       ld_once<float>(f3, po, cb, b4));
  }
  // stores r and r*v; returns the velocity the CFL reduction sees (stored m / stored rho)
  template <class AR>
  static __device__ __forceinline__ float3 store(void* o0, void* o1, void* o2, void* o3, long long po, unsigned cb,
                                                 unsigned b4, float r, typename AR::Rcp rr, float v0, float v1,
                                                 float v2, bool want_v) {
    const float m0 = r * v0, m1 = r * v1, m2 = r * v2;
    st_once<float>(o0, po, cb, b4, r);
    st_once<float>(o1, po, cb, b4, m0);
    st_once<float>(o2, po, cb, b4, m1);
    st_once<float>(o3, po, cb, b4, m2);
    if (!want_v) return make_float3(0.0f, 0.0f, 0.0f);
    const float m[3] = {m0, m1, m2};
    float v[3];
    AR::template divn<3>(m, rr, v);
    return make_float3(v[0], v[1], v[2]);
  }
};

struct F16V {
  static constexpr bool VEL = true;    // stores velocity
  static __device__ __forceinline__ float4 load(const void* f0, const void* f1, const void* f2, const void* f3,
                                                long long po, unsigned cb, unsigned b4) {
    return make_float4(ldp<float>(f0, po, cb, b4), h2f(ldp<uint16_t>(f1, po, cb, b4)),
                       h2f(ldp<uint16_t>(f2, po, cb, b4)), h2f(ldp<uint16_t>(f3, po, cb, b4)));
  }
  static __device__ __forceinline__ float4 load_once(const void* f0, const void* f1, const void* f2, const void* f3,
                                                     long long po, unsigned cb, unsigned b4) {
    return make_float4(ld_once<float>(f0, po, cb, b4), h2f(ld_once<uint16_t>(f1, po, cb, b4)),
                       h2f(ld_once<uint16_t>(f2, po, cb, b4)), h2f(ld_once<uint16_t>(f3, po, cb, b4)));
  }
  template <class AR>
  static __device__ __forceinline__ float3 store(void* o0, void* o1, void* o2, void* o3, long long po, unsigned cb,
                                                 unsigned b4, float r, typename AR::Rcp, float v0, float v1, float v2,
                                                 bool) {
    const uint16_t h0 = f32_to_f16_bits(v0), h1 = f32_to_f16_bits(v1), h2 = f32_to_f16_bits(v2);
    st_once<float>(o0, po, cb, b4, r);
    st_once<uint16_t>(o1, po, cb, b4, h0);
    st_once<uint16_t>(o2, po, cb, b4, h1);
    st_once<uint16_t>(o3, po, cb, b4, h2);
    return make_float3(h2f(h0), h2f(h1), h2f(h2));   // the velocity the next step will read
  }
};

// primitive state (nebula3d.prim, nebula3d.py:28-32) and conserved state of a stored cell
template <class ST, class AR>
__device__ __forceinline__ void decode(const float4& raw, float floor_, float4& q, float4& U) {
  if constexpr (ST::VEL) {
    q = make_float4(fmaxf(raw.x, floor_), raw.y, raw.z, raw.w);
    U = make_float4(raw.x, raw.x * raw.y, raw.x * raw.z, raw.x * raw.w);
  } else {
    const float r = fmaxf(raw.x, floor_);
    const typename AR::Rcp rr = AR::rcp(r);
    const float m[3] = {raw.y, raw.z, raw.w};
    float v[3];
    AR::template divn<3>(m, rr, v);
    q = make_float4(r, v[0], v[1], v[2]);
    U = raw;
  }
}

template <class ST>
__device__ __forceinline__ float4 conserved(const float4& raw) {
  if constexpr (ST::VEL) return make_float4(raw.x, raw.x * raw.y, raw.x * raw.z, raw.x * raw.w);
  else return raw;
}

// ------------------------------------------------------------------------------------------ physics
// slope limiters: minmod (nebula3d.py:40-46), van Leer (sim.py:375-380), MC (sim.py:366-372)
template <class AR, int LIM>
__device__ __forceinline__ float slope(float a, float b) {
  if constexpr (AR::FAST && LIM == NV_LIM_MINMOD) {
    // the median of (a, b, 0) is minmod(a, b): 0 for opposite signs or a zero operand, else the operand nearer
    // zero. It differs from the product test only where a*b underflows to zero (|slope| < 2^-75).
    return __builtin_amdgcn_fmed3f(a, b, 0.0f);
  }
  const float p = a * b;
  if constexpr (LIM == NV_LIM_VANLEER) {
    float v;
    if constexpr (AR::FAST) v = AR::div(2.0f * a * b, AR::rcp(a + b));   // a + b = 0 only where p < 0
    else v = 2.0f * a * b / (a + b);
    return p <= 0.0f ? 0.0f : v;
  } else if constexpr (LIM == NV_LIM_MC) {
    const float s = a > 0.0f ? 1.0f : -1.0f;
    const float v = s * fminf(fminf(2.0f * fabsf(a), 2.0f * fabsf(b)), 0.5f * fabsf(a + b));
    return p <= 0.0f ? 0.0f : v;
  } else {
    // a*b > 0: a and b are nonzero with one sign, and the median of (a, b, 0) is the one nearer zero, which is
    // the reference's choice (ties |a| == |b| mean a == b)
    const float m = __builtin_amdgcn_fmed3f(a, b, 0.0f);
    return p <= 0.0f ? 0.0f : m;
  }
}

__device__ __forceinline__ float comp(const float4& q, int c) {
  return c == 0 ? q.x : (c == 1 ? q.y : (c == 2 ? q.z : q.w));
}

// Edge states of a cell from its own and its two neighbours' primitive states along one axis: the limited
// slope s = lim(q - q-, q+ - q) is the same for the cell's low and high faces, so lo = q - s/2 is the right
// state R of the face below (nebula3d.py:61, sim.py:402) and hi = q + s/2 the left state L of the face above
// (nebula3d.py:60, sim.py:401). In the exact policy mad(-1/2, s, v) is (-s/2) + v, i.e. v - s/2 bit for bit.
template <class AR, int LM>
__device__ __forceinline__ void edges(const float4& qm, const float4& q, const float4& qp, float4& lo, float4& hi) {
  float l[4], h[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const float v = comp(q, c);
    const float sl = slope<AR, LM>(v - comp(qm, c), comp(qp, c) - v);
    l[c] = AR::mad(-0.5f, sl, v);
    h[c] = AR::mad(0.5f, sl, v);
  }
  lo = make_float4(l[0], l[1], l[2], l[3]);
  hi = make_float4(h[0], h[1], h[2], h[3]);
}

// Face flux along axis AX (0=x, 1=y, 2=z) from the reconstructed states L (high edge of the cell below) and R
// (low edge of the cell above), with c^2 and c of those two cells. HLL (sim.py:397-420) or Rusanov
// (nebula3d.py:54-72). The reference's one-hot axis vector e is folded in at compile time: dot(v, e) = v[AX+1],
// e*(r c^2) touches component AX+1 only. Sums are written as mad(x, y, z) = x*y + z, which in the exact
// policy is the reference expression with its two addends swapped (the same IEEE result).
template <class AR, int AX, int SOLVER>
__device__ __forceinline__ float4 face_flux(const float4& Lq, const float4& Rq, float c2b, float c2c, float cb,
                                            float cc, float floor_) {
  const float L[4] = {Lq.x, Lq.y, Lq.z, Lq.w};
  const float R[4] = {Rq.x, Rq.y, Rq.z, Rq.w};
  const float rL = fmaxf(L[0], floor_), rR = fmaxf(R[0], floor_);
  const float unL = L[AX + 1], unR = R[AX + 1];
  const float UL[4] = {rL, rL * L[1], rL * L[2], rL * L[3]};
  const float UR[4] = {rR, rR * R[1], rR * R[2], rR * R[3]};
  float FL[4], FR[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    FL[c] = UL[c] * unL;
    FR[c] = UR[c] * unR;
  }
  FL[AX + 1] = AR::mad(rL, c2b, FL[AX + 1]);
  FR[AX + 1] = AR::mad(rR, c2c, FR[AX + 1]);
  float F[4];
  if constexpr (SOLVER == NV_SOLVER_RUSANOV) {
    const float amax = fmaxf(fabsf(unL) + cb, fabsf(unR) + cc);
    const float ha = 0.5f * amax;
#pragma unroll
    for (int c = 0; c < 4; ++c) F[c] = AR::mad(-ha, UR[c] - UL[c], 0.5f * (FL[c] + FR[c]));
  } else if constexpr (AR::FAST) {
    // The HLL flux (SR FL - SL FR + SL SR (UR - UL)) / (SR - SL) as wR FL + wL FR + s (UR - UL) with
    // wR = SR / (SR - SL), wL = -SL / (SR - SL), s = SL SR / (SR - SL); since FL = UL unL (+ rL c^2 on the normal
    // component) and UL = rL (1, vL), every component is F_c = L_c bL + R_c bR (L_0 = R_0 = 1) with
    // bL = rL (wR unL - s) >= 0 and bR = rR (wL unR + s) <= 0, plus wR rL c^2_b + wL rR c^2_c on the normal one.
    // SL clamped to <= 0 and SR to >= 0 make the one formula give FL where SL >= 0 (wR = 1, wL = s = 0) and FR
    // where SR <= 0, with no three-way select; SR - SL > 0 since c > 0.
    const float SL = fminf(fminf(unL - cb, unR - cc), 0.0f);
    const float SR = fmaxf(fmaxf(unL + cb, unR + cc), 0.0f);
    const float inv = AR::rcp(fmaxf(SR - SL, 1.17549435e-38f));
    const float wR = SR * inv, wL = -SL * inv;
    const float sd = SL * wR;
    const float bL = rL * AR::mad(wR, unL, -sd), bR = rR * AR::mad(wL, unR, sd);
    F[0] = bL + bR;
#pragma unroll
    for (int c = 1; c < 4; ++c) F[c] = AR::mad(L[c], bL, R[c] * bR);
    F[AX + 1] = AR::mad(wR * rL, c2b, AR::mad(wL * rR, c2c, F[AX + 1]));
  } else {
    const float SL = fminf(unL - cb, unR - cc);
    const float SR = fmaxf(unL + cb, unR + cc);
    if (SL >= 0.0f) {
#pragma unroll
      for (int c = 0; c < 4; ++c) F[c] = FL[c];
    } else if (SR <= 0.0f) {
#pragma unroll
      for (int c = 0; c < 4; ++c) F[c] = FR[c];
    } else {
      const float ss = SL * SR;
      const typename AR::Rcp rden = AR::rcp(SR - SL);
      float N[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) N[c] = AR::mad(ss, UR[c] - UL[c], AR::mad(SR, FL[c], -(SL * FR[c])));
      AR::template divn<4>(N, rden, F);
    }
  }
  return make_float4(F[0], F[1], F[2], F[3]);
}

// One plane of a cell column's register window: primitives, conserved state, xi, and c^2 and c.
struct Win {
  float4 q, U;
  float x, c2, c;
};

// Which stage forms load the halo of the next plane one plane ahead (four more registers per thread): stage 1; and
// with QUEUE (the batch's persistent tile-queue kernel), stage 2 in the fast float32 form too (123 registers). The
// exact and binary16-velocity forms' stage 2, which also hold U0 and the CFL max, would spill at their 128-register
// limit; the single-grid fast stage 2 fits but measured slower with it (1.52 -> 1.60 ms at 512^3), the batch's
// queue kernel faster (1.07 -> 1.02 ms per launch).
template <bool STAGE1, bool QUEUE, class ST, class AR>
struct HALO_PREFETCH {
  static constexpr bool value = STAGE1 || (QUEUE && AR::FAST && !ST::VEL);
};

// One tile (64 z x 8 y columns, marching along x over ci planes) of one RK stage.
// STAGE1: a = 0, b = 1 and U0 = U (sim.py:1025), so out = 0*U0 + 1*Un is Un (up to the sign of zero).
template <class TL, class ST, class AR, int SOLVER, int LIM, bool SMAX, bool STAGE1, bool QUEUE = false>
__device__ __forceinline__ void stage_tile(const HydroArgs& A, int tile) {
  constexpr int TK = TL::TK, TJ = TL::TJ, PK = TL::PK, PJ = TL::PJ, CK = TL::CK, CJ = TL::CJ;
  constexpr int NHALO = TL::NHALO, NCHALO = TL::NCHALO, NT = TL::NT;
  constexpr int LM = SOLVER == NV_SOLVER_HLL ? LIM : NV_LIM_MINMOD;   // Rusanov always uses minmod
  using Rcp = typename AR::Rcp;
  __shared__ float4 sq[PJ * PK];            // primitives of plane i, 2-cell halo
  __shared__ float2 sc[CJ * CK];            // (c^2, c) of plane i, 1-cell halo
  __shared__ float4 ehy[(TJ + 1) * TK];     // high y edge of rows -1 .. TJ-1 (row r+1 holds row r)
  __shared__ float4 ehz[TJ * (TK + 1)];     // high z edge of columns -1 .. TK-1
  __shared__ float4 ely[TK];                // low y edge of row TJ (the next tile's first row)
  __shared__ float4 elz[TJ];                // low z edge of column TK
  __shared__ float4 fy[(TJ + 1) * TK];      // y-face fluxes: row r is the face between rows r-1 and r
  __shared__ float4 fz[TJ * (TK + 1)];      // z-face fluxes: column t is the face between columns t-1 and t
  __shared__ float red[NT / 64];
  __shared__ int2 htab[NHALO];              // primitive-halo cell h: (LDS slot, clamped offset within a plane)
  __shared__ int2 ctab[NCHALO];             // c^2-halo cell h: the same

  const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * TK + tx;
  const int nx = A.nx, ny = A.ny, nz = A.nz;
  const int bk = tile % A.tiles_k, bj = (tile / A.tiles_k) % A.tiles_j, bi = tile / (A.tiles_k * A.tiles_j);
  const int k0 = bk * TK, j0 = bj * TJ, i0 = bi * A.ci;
  const int i1 = min(i0 + A.ci, nx);
  const int k = k0 + tx, j = j0 + ty;
  const bool active = (k < nz) && (j < ny);
  const long long plane = (long long)ny * nz;
  // own column, clamped (threads past the domain edge still supply the clamped ghost values of ldc)
  const unsigned cb = 4u * (unsigned)(min(j, ny - 1) * nz + min(k, nz - 1));   // its byte offset in a plane
  const unsigned pb4 = 4u * (unsigned)plane;                                     // bytes of a float plane
  const float floor_ = A.floor_, cn2 = A.cn2, dci = A.ci2 - A.cn2;
  const float dt = A.dt_dev ? *A.dt_dev : A.dt;
  // The switched-on stars (weight > 0) and, among them, those with an active jet (strength > 0), one bit per star:
  // lane l of each wavefront looks at star l once per tile, so a cell loops over these stars only, in the same
  // ascending order, instead of loading and testing every star's table entries (stars that fail the test add
  // nothing). Above 64 stars the cells test every star themselves.
  bool on_l = false, jet_l = false;
  {
    const int l = __lane_id();
    if (l < A.nstars) {
      on_l = A.stars[l].w > 0.0f;
      jet_l = on_l && A.jets[l].w > 0.0f;
    }
  }
  const unsigned long long star_bits = __ballot(on_l), jet_bits = __ballot(jet_l);
  const bool all_stars = A.nstars > 64;
  auto each_star = [&](unsigned long long bits, auto&& fn) {
    if (all_stars) {
      for (int s = 0; s < A.nstars; ++s) fn(s);
    } else {
      for (unsigned long long m = bits; m; m &= m - 1) fn(__builtin_ctzll(m));
    }
  };

  auto load_col = [&](int p, float4& raw, float& x) {
    const long long po = (long long)clampi(p, 0, nx - 1) * plane;
    raw = ST::load(A.f0, A.f1, A.f2, A.f3, po, cb, pb4);
    x = ldp<float>(A.xi, po, cb, pb4);
  };

  // the halo cells' LDS slots and in-plane offsets depend on the tile only: tabulated once for the whole march
  for (int h = tid; h < NHALO; h += NT) {
    int jj, kk;
    if (h < 2 * PK) {
      jj = h / PK;
      kk = h - jj * PK;
    } else if (h < 4 * PK) {
      const int hh = h - 2 * PK;
      jj = TJ + 2 + hh / PK;
      kk = hh - (hh / PK) * PK;
    } else {
      const int hh = h - 4 * PK;
      jj = 2 + (hh >> 2);
      const int c = hh & 3;
      kk = c < 2 ? c : TK + c;
    }
    htab[h] = make_int2(jj * PK + kk, clampi(j0 + jj - 2, 0, ny - 1) * nz + clampi(k0 + kk - 2, 0, nz - 1));
  }
  for (int h = tid; h < NCHALO; h += NT) {
    int jj, kk;
    if (h < 2 * CK) {
      jj = h < CK ? 0 : TJ + 1;
      kk = h < CK ? h : h - CK;
    } else {
      const int hh = h - 2 * CK;
      jj = 1 + (hh >> 1);
      kk = (hh & 1) ? TK + 1 : 0;
    }
    ctab[h] = make_int2(jj * CK + kk, clampi(j0 + jj - 1, 0, ny - 1) * nz + clampi(k0 + kk - 1, 0, nz - 1));
  }
  __syncthreads();

  // register window of the own column: three slots holding planes i, i+1, i+2 (primitives, conserved state, xi;
  // c^2 and c of the first two), plus the high x edge of plane i and the flux of face i-1/2.
  Win w0, w1, w2;
  float4 hix, fLx;
  {
    float4 raw, q0, q1, Uc, lo, hm1;
    float x, c2m1, cm1;
    load_col(i0 - 2, raw, x);
    decode<ST, AR>(raw, floor_, q0, Uc);
    load_col(i0 - 1, raw, x);
    decode<ST, AR>(raw, floor_, q1, Uc);
    c2m1 = AR::mad(x, dci, cn2);
    cm1 = AR::sqrt(c2m1);
    load_col(i0, raw, w0.x);
    decode<ST, AR>(raw, floor_, w0.q, w0.U);
    w0.c2 = AR::mad(w0.x, dci, cn2);
    w0.c = AR::sqrt(w0.c2);
    load_col(i0 + 1, raw, w1.x);
    decode<ST, AR>(raw, floor_, w1.q, w1.U);
    w1.c2 = AR::mad(w1.x, dci, cn2);
    w1.c = AR::sqrt(w1.c2);
    load_col(i0 + 2, raw, w2.x);
    decode<ST, AR>(raw, floor_, w2.q, w2.U);
    edges<AR, LM>(q0, q1, w0.q, lo, hm1);     // plane i0-1
    edges<AR, LM>(q1, w0.q, w1.q, lo, hix);   // plane i0
    fLx = face_flux<AR, 0, SOLVER>(hm1, lo, c2m1, w0.c2, cm1, w0.c, floor_);
  }

  const Rcp rdx = AR::rcp(A.dx);
  const float dx = A.dx;
  float tmax = 0.0f;

  // The halo of plane p (at most one cell per thread: the primitive halo on the last NHALO threads, the c^2 halo
  // on the first NCHALO): with PF it is loaded into registers one plane ahead, so its latency is not exposed
  // between the start of a plane and its first barrier.
  static_assert(NHALO <= NT && NCHALO <= NT && NCHALO <= NT - NHALO, "one halo cell per thread");
  constexpr bool PF = HALO_PREFETCH<STAGE1, QUEUE, ST, AR>::value;
  const bool hq = tid >= NT - NHALO, hc = tid < NCHALO;
  auto halo_load = [&](int p, float4& raw) {
    const long long pb = (long long)p * plane;
    if (hq) raw = ST::load(A.f0, A.f1, A.f2, A.f3, pb, 4u * (unsigned)htab[NT - 1 - tid].y, pb4);
    else if (hc) raw.x = ldp<float>(A.xi, pb, 4u * (unsigned)ctab[tid].y, pb4);
  };
  float4 hraw = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if constexpr (PF) halo_load(i0, hraw);

  // plane i with the window slots A, B, C holding planes i, i+1, i+2; plane i+3 is loaded into slot A
  auto plane_step = [&](Win& A_, Win& B_, Win& C_, int i) __attribute__((always_inline)) {
    // next window plane (i+3): issued now, consumed at the end of this iteration
    float4 rnext;
    float xnext;
    load_col(i + 3, rnext, xnext);
    float4 U0raw = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const long long po = (long long)i * plane;
    if (!STAGE1 && !A.u0_is_u && active) U0raw = ST::load_once(A.g0, A.g1, A.g2, A.g3, po, cb, pb4);

    // ---- A: plane i into LDS (own cell from the window, halo from global)
    const int cq = (ty + 2) * PK + tx + 2;
    const int cc = (ty + 1) * CK + tx + 1;
    sq[cq] = A_.q;
    sc[cc] = make_float2(A_.c2, A_.c);
    // halo cells from the last wavefronts up (the first ones take the c^2 halo here and the tile-boundary edges
    // and faces later), so that no wavefront carries two of these jobs between two barriers
    if constexpr (PF) {
      if (hq) {
        float4 qh, Uh;
        decode<ST, AR>(hraw, floor_, qh, Uh);
        sq[htab[NT - 1 - tid].x] = qh;
      } else if (hc) {
        const float c2 = AR::mad(hraw.x, dci, cn2);
        sc[ctab[tid].x] = make_float2(c2, AR::sqrt(c2));
      }
    } else {
      const long long pbase = (long long)i * plane;
      for (int h = NT - 1 - tid; h < NHALO; h += NT) {
        const int2 e = htab[h];
        float4 qh, Uh;
        decode<ST, AR>(ST::load(A.f0, A.f1, A.f2, A.f3, pbase, 4u * (unsigned)e.y, pb4), floor_, qh, Uh);
        sq[e.x] = qh;
      }
      for (int h = tid; h < NCHALO; h += NT) {
        const int2 e = ctab[h];
        const float c2 = AR::mad(ldp<float>(A.xi, pbase, 4u * (unsigned)e.y, pb4), dci, cn2);
        sc[e.x] = make_float2(c2, AR::sqrt(c2));
      }
    }
    __syncthreads();
    if (PF && i + 1 < i1) halo_load(i + 1, hraw);

    // ---- B: edge states. Own cell in y and z; the tile's outer neighbours for the boundary faces.
    float4 loy, loz, hi;
    edges<AR, LM>(sq[cq - PK], A_.q, sq[cq + PK], loy, hi);   // the own cell is A_.q, which sq[cq] holds
    ehy[(ty + 1) * TK + tx] = hi;
    edges<AR, LM>(sq[cq - 1], A_.q, sq[cq + 1], loz, hi);
    ehz[ty * (TK + 1) + tx + 1] = hi;
    {
      float4 lo_, hi_;
      if (tid < TK) {                        // row -1: high y edge
        const int bq = PK + tid + 2;
        edges<AR, LM>(sq[bq - PK], sq[bq], sq[bq + PK], lo_, hi_);
        ehy[tid] = hi_;
      } else if (tid < 2 * TK) {             // row TJ: low y edge
        const int t = tid - TK;
        const int bq = (TJ + 2) * PK + t + 2;
        edges<AR, LM>(sq[bq - PK], sq[bq], sq[bq + PK], lo_, hi_);
        ely[t] = lo_;
      } else if (tid < 2 * TK + TJ) {        // column -1: high z edge
        const int r = tid - 2 * TK;
        const int bq = (r + 2) * PK + 1;
        edges<AR, LM>(sq[bq - 1], sq[bq], sq[bq + 1], lo_, hi_);
        ehz[r * (TK + 1)] = hi_;
      } else if (tid < 2 * TK + 2 * TJ) {    // column TK: low z edge
        const int r = tid - 2 * TK - TJ;
        const int bq = (r + 2) * PK + TK + 2;
        edges<AR, LM>(sq[bq - 1], sq[bq], sq[bq + 1], lo_, hi_);
        elz[r] = lo_;
      }
    }
    // x: the window gives plane i+1's edges; its low edge closes face i+1/2, its high edge is carried
    float4 lox1, hix1;
    edges<AR, LM>(A_.q, B_.q, C_.q, lox1, hix1);
    const float4 fRx = face_flux<AR, 0, SOLVER>(hix, lox1, A_.c2, B_.c2, A_.c, B_.c, floor_);
    __syncthreads();

    // ---- C: the low y and z faces of every cell, and the tile's high y and z boundary faces
    {
      const float2 cy0 = sc[cc - CK], c1 = sc[cc], cz0 = sc[cc - 1];
      fy[ty * TK + tx] = face_flux<AR, 1, SOLVER>(ehy[ty * TK + tx], loy, cy0.x, c1.x, cy0.y, c1.y, floor_);
      fz[ty * (TK + 1) + tx] = face_flux<AR, 2, SOLVER>(ehz[ty * (TK + 1) + tx], loz, cz0.x, c1.x, cz0.y, c1.y,
                                                        floor_);
      if (tid < TK) {
        const int bc = TJ * CK + tid + 1;    // c plane: row TJ-1 at bc, row TJ at bc + CK
        const float2 a0 = sc[bc], a1 = sc[bc + CK];
        fy[TJ * TK + tid] = face_flux<AR, 1, SOLVER>(ehy[TJ * TK + tid], ely[tid], a0.x, a1.x, a0.y, a1.y, floor_);
      } else if (tid < TK + TJ) {
        const int r = tid - TK;
        const int bc = (r + 1) * CK + TK;    // column TK-1 at bc, column TK at bc + 1
        const float2 a0 = sc[bc], a1 = sc[bc + 1];
        fz[r * (TK + 1) + TK] = face_flux<AR, 2, SOLVER>(ehz[r * (TK + 1) + TK], elz[r], a0.x, a1.x, a0.y, a1.y,
                                                         floor_);
      }
    }
    __syncthreads();

    // ---- D: update of the own cell
    if (active) {
      const float4 fLy = fy[ty * TK + tx], fRy = fy[(ty + 1) * TK + tx];
      const float4 fLz = fz[ty * (TK + 1) + tx], fRz = fz[ty * (TK + 1) + tx + 1];
      float div[4];
#pragma unroll
      for (int c = 0; c < 4; ++c)   // the reference's 0 + ... only turns a -0 sum into +0: dropped
        div[c] = ((comp(fRx, c) - comp(fLx, c)) + (comp(fRy, c) - comp(fLy, c))) + (comp(fRz, c) - comp(fLz, c));
      const float4 Uc = A_.U;
      const float U[4] = {Uc.x, Uc.y, Uc.z, Uc.w};
      // T = Un (sim.py:486) being formed: the fast policy starts from U - (dt / dx) div with one scale and adds
      // dt rho a for each source; the exact one forms dU = -div / dx plus rho a per source, then U + dt dU
      float T[4];
      if constexpr (AR::FAST) {
        const float kdt = -dt * (A.inv_dx != 0.0f ? A.inv_dx : rdx);
#pragma unroll
        for (int c = 0; c < 4; ++c) T[c] = AR::mad(div[c], kdt, U[c]);
      } else if (A.inv_dx != 0.0f) {
        // dx = 2^e: -div * 2^-e is the same real number as -div / dx, so one rounding gives the same float
#pragma unroll
        for (int c = 0; c < 4; ++c) T[c] = -div[c] * A.inv_dx;
      } else {
        const float nd[4] = {-div[0], -div[1], -div[2], -div[3]};
        AR::template divn<4>(nd, rdx, T);
      }
      auto add_source = [&](float a0, float a1, float a2) {   // a momentum source rho (a0, a1, a2)
        if constexpr (AR::FAST) {
          const float dtr = dt * U[0];
          T[1] = AR::mad(dtr, a0, T[1]);
          T[2] = AR::mad(dtr, a1, T[2]);
          T[3] = AR::mad(dtr, a2, T[3]);
        } else {
          T[1] = T[1] + U[0] * a0;
          T[2] = T[2] + U[0] * a1;
          T[3] = T[3] + U[0] * a2;
        }
      };
      const float p0 = ((float)i + 0.5f) * dx, p1 = ((float)j + 0.5f) * dx, p2 = ((float)k + 0.5f) * dx;
      const float x_ion = A_.x;
      if (A.wind > 0.0f && x_ion > 0.01f) {
        // radiation pressure / wind on ionized gas (sim.py:463-473)
        float acc0 = 0.0f, acc1 = 0.0f, acc2 = 0.0f;
        each_star(star_bits, [&](int s) {
          const float4 st = A.stars[s];
          if (st.w > 0.0f) {
            const float r0 = p0 - st.x, r1 = p1 - st.y, r2v = p2 - st.z;
            const float r2 = ((r0 * r0 + r1 * r1) + r2v * r2v) + 0.0004f;
            float f;
            if constexpr (AR::FAST) {
              const float ir = __builtin_amdgcn_rsqf(r2);   // st.w / r2^(3/2) to a few ulp
              f = st.w * (ir * ir * ir);
            } else {
              f = Exact::div(st.w, recip_f(r2 * sqrt_rn(r2)));
            }
            acc0 = acc0 + r0 * f;
            acc1 = acc1 + r1 * f;
            acc2 = acc2 + r2v * f;
          }
        });
        const float wx = A.wind * x_ion;
        add_source(acc0 * wx, acc1 * wx, acc2 * wx);
      }
      // bipolar jets (sim.py:474-485); where no cone contributes the term is +0 and adding it is skipped
      float aj0 = 0.0f, aj1 = 0.0f, aj2 = 0.0f;
      bool jet_hit = false;
      each_star(jet_bits, [&](int s) {
        const float4 jt = A.jets[s];
        const float4 st = A.stars[s];
        if (jt.w > 0.0f && st.w > 0.0f) {
          const float r0 = p0 - st.x, r1 = p1 - st.y, r2v = p2 - st.z;
          const float d2 = (r0 * r0 + r1 * r1) + r2v * r2v;
          // sqrt(d2) >= jet_r wherever d2 > jet_r2_hi: outside the cone's ball without a square root
          if (d2 <= A.jet_r2_hi) {
            const float r = sqrt_rn(d2);
            if (r < A.jet_r && r > 1e-6f) {
              const float cth = ((r0 * jt.x + r1 * jt.y) + r2v * jt.z) / r;
              if (fabsf(cth) > A.jet_cos) {
                const float f = (cth > 0.0f ? 1.0f : -1.0f) * jt.w * A.pulse / (r * r + 0.0004f);
                aj0 = aj0 + jt.x * f;
                aj1 = aj1 + jt.y * f;
                aj2 = aj2 + jt.z * f;
                jet_hit = true;
              }
            }
          }
        }
      });
      if (jet_hit) add_source(aj0, aj1, aj2);
      float Un[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) Un[c] = AR::FAST ? T[c] : AR::mad(dt, T[c], U[c]);

      float out[4];
      if constexpr (STAGE1) {
#pragma unroll
        for (int c = 0; c < 4; ++c) out[c] = Un[c];
      } else {
        const float4 U0c = conserved<ST>(U0raw);
        const float U0v[4] = {A.u0_is_u ? U[0] : U0c.x, A.u0_is_u ? U[1] : U0c.y, A.u0_is_u ? U[2] : U0c.z,
                              A.u0_is_u ? U[3] : U0c.w};
#pragma unroll
        for (int c = 0; c < 4; ++c) out[c] = AR::mad(A.a, U0v[c], A.b * Un[c]);
      }
      const float r = fmaxf(out[0], floor_);
      if constexpr (AR::FAST && !ST::VEL) {
        // The stored momentum is r (m / r) = m, so m is stored as it is unless |m|^2 may exceed (vmax r)^2; the
        // threshold sits a little below vmax^2 so that every cell the cap can touch takes the exact test.
        float m0 = out[1], m1 = out[2], m2 = out[3];
        const float s2m = AR::mad(m2, m2, AR::mad(m1, m1, m0 * m0));
        if (s2m > (A.vmax2_lo * 0.99999f) * (r * r)) {
          const Rcp rc = AR::rcp(r);
          const float v0 = m0 * rc, v1 = m1 * rc, v2 = m2 * rc;
          const float sp = AR::sqrt(AR::mad(v2, v2, AR::mad(v1, v1, v0 * v0)));
          const float f = sp > A.vmax ? A.vmax * AR::rcp(sp) : 1.0f;   // the cap of sim.py:491-494
          m0 = r * (v0 * f);
          m1 = r * (v1 * f);
          m2 = r * (v2 * f);
        }
        st_once<float>(A.o0, po, cb, pb4, r);
        st_once<float>(A.o1, po, cb, pb4, m0);
        st_once<float>(A.o2, po, cb, pb4, m1);
        st_once<float>(A.o3, po, cb, pb4, m2);
        if constexpr (SMAX) tmax = fmaxf(tmax, ((fabsf(m0) + fabsf(m1)) + fabsf(m2)) * AR::rcp(r) + A_.c);
      } else {
      const Rcp rr = AR::rcp(r);
      const float mo[3] = {out[1], out[2], out[3]};
      float vo[3];
      AR::template divn<3>(mo, rr, vo);
      float v0 = vo[0], v1 = vo[1], v2 = vo[2];
      const float s2 = AR::mad(v2, v2, AR::mad(v1, v1, v0 * v0));
      if (s2 > A.vmax2_lo) {
        // |v| = sqrt(s2) may exceed vmax: the reference test and cap (sim.py:491-494)
        const float sp = sqrtf(s2);
        if (sp > A.vmax) {
          const float f = A.vmax / sp;
          v0 = v0 * f;
          v1 = v1 * f;
          v2 = v2 * f;
        }
      }
      const float3 vs = ST::template store<AR>(A.o0, A.o1, A.o2, A.o3, po, cb, pb4, r, rr, v0, v1, v2, SMAX);
      if constexpr (SMAX) {
        // max_speed of the written state (nebula3d.py:154-162): L1 norm of its velocity plus this cell's
        // sound speed sqrt(cn2 + xi (ci2 - cn2)), which is A_.c
        const float s = ((fabsf(vs.x) + fabsf(vs.y)) + fabsf(vs.z)) + A_.c;
        tmax = fmaxf(tmax, s);
      }
      }
    }
    fLx = fRx;
    hix = hix1;

    // ---- advance the window by one plane: plane i+2 (slot C) gets its c^2 and c, plane i+3 takes slot A
    C_.c2 = AR::mad(C_.x, dci, cn2);
    C_.c = AR::sqrt(C_.c2);
    decode<ST, AR>(rnext, floor_, A_.q, A_.U);
    A_.x = xnext;
  };
  // The compiler turns the rotation into register moves (37 per plane). Unrolling the march three ways, one
  // rotation per copy, measured 4-7% slower per launch at 512^3; with scheduling barriers between the copies 17
  // moves per plane remain, stage 2 ran 3% faster and stage 1 16% slower, and the file took three times as long to
  // compile.
  for (int i = i0; i < i1; ++i) {
    plane_step(w0, w1, w2, i);
    const Win t = w0;
    w0 = w1;
    w1 = w2;
    w2 = t;
  }
  if constexpr (SMAX) block_max_atomic(tmax, red, A.smax);
}

// Each XCD takes runs of consecutive tiles (z fastest, then y, then the x chunk, then the next member of a batch),
// so the tiles whose halos overlap run on one XCD and share its L2. With run = per_xcd an XCD takes one contiguous
// run (xcd_tile). When the tiles outnumber the workgroups the device holds at once, run is the number an XCD holds:
// the XCDs then take bands of run tiles side by side, band after band, so the workgroups resident at one time
// cover 8 run consecutive tiles, and a band's neighbour rows are being read at the same time by the next XCD
// (through the shared memory-side cache) rather than a whole round of workgroups earlier or later. The dispatcher
// starts an XCD's workgroups in launch order, so its s-th workgroup takes tile (s / run) 8 run + x run + s % run.
__device__ __forceinline__ int band_tile(int per_xcd, int run) {
  const int x = blockIdx.x & 7, s = blockIdx.x >> 3;
  return run >= per_xcd ? x * per_xcd + s : (s / run) * (8 * run) + x * run + (s - (s / run) * run);
}

template <class TL, class ST, class AR, int SOLVER, int LIM, bool SMAX, bool STAGE1>
__global__ __launch_bounds__(TL::NT, 4) void hydro_stage_kernel(HydroArgs A) {
  const int tile = band_tile(A.per_xcd, A.run);
  if (tile >= A.tiles) return;
  stage_tile<TL, ST, AR, SOLVER, LIM, SMAX, STAGE1>(A, tile);
}

// A batch of independent grids in one launch: tile t of the flat list belongs to the member s with
// start[s] <= t < start[s+1] (start[0] = 0, start[n] = total), whose arguments are args[s].
__device__ __forceinline__ int member_of(const int* __restrict__ start, int n, int t) {
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (start[mid] <= t) lo = mid;
    else hi = mid;
  }
  return lo;
}

// One workgroup per tile, dealt by the hardware dispatcher.
template <class TL, class ST, class AR, int SOLVER, int LIM, bool SMAX, bool STAGE1>
__global__ __launch_bounds__(TL::NT, 4) void hydro_stage_batch_kernel(const HydroArgs* __restrict__ args,
                                                                  const int* __restrict__ start, int n, int total,
                                                                  int per_xcd) {
  const int t = xcd_tile(per_xcd);
  if (t >= total) return;
  const int lo = member_of(start, n, t);
  stage_tile<TL, ST, AR, SOLVER, LIM, SMAX, STAGE1>(args[lo], t - start[lo]);
}

// The persistent form: a grid of resident workgroups fed by a tile work queue. queue[x], x < 8, counts the tiles
// handed out of XCD x's contiguous run [x per_xcd, (x + 1) per_xcd); a workgroup claims from its own XCD's run first
// and then steals from the others in turn, so tiles stay XCD-local while they last and no CU idles while another
// XCD still has work. Every claim is one atomic add, so each tile is claimed once, and a workgroup exits after
// finding all eight runs exhausted (at most eight adds per claim). queue[8] counts the workgroups that have exited;
// the last one resets all nine counters for the next launch of the plan (launches of one plan must not overlap).
// Looping over tiles keeps more values live than one tile per workgroup does (the compiler hoists the tile body's
// invariants), so this form spills a few registers; which form is faster is measured (SimBatch dispatch="tune").
template <class TL, class ST, class AR, int SOLVER, int LIM, bool SMAX, bool STAGE1>
__global__ __launch_bounds__(TL::NT, 4) void hydro_stage_queue_kernel(const HydroArgs* __restrict__ args,
                                                                  const int* __restrict__ start, int n, int total,
                                                                  int per_xcd, int* __restrict__ queue) {
  __shared__ int claimed;
  const bool lead = threadIdx.x == 0 && threadIdx.y == 0;
  const int home = blockIdx.x & 7;
  for (;;) {
    if (lead) {
      int t = -1;
      for (int q = 0; q < 8 && t < 0; ++q) {
        const int x = (home + q) & 7;
        const int base = x * per_xcd, cnt = min(per_xcd, total - base);
        if (cnt <= 0) continue;
        const int c = atomicAdd(&queue[x], 1);
        if (c < cnt) t = base + c;
      }
      claimed = t;
    }
    __syncthreads();
    const int t = claimed;
    __syncthreads();                  // every thread has read the claim before the lead writes the next one
    if (t < 0) break;
    const int lo = member_of(start, n, t);
    stage_tile<TL, ST, AR, SOLVER, LIM, SMAX, STAGE1, true>(args[lo], t - start[lo]);
  }
  if (lead) {
    __threadfence();
    if (atomicAdd(&queue[8], 1) == (int)gridDim.x - 1) {
      for (int x = 0; x < 9; ++x) atomicExch(&queue[x], 0);
      __threadfence();
    }
  }
}

// ------------------------------------------------------------------------------------------ conservation sums
// Mass budget of a stage in FP64: out[0] += sum rho, out[1] += sum over the six domain faces of the outward mass
// flux of the stage's face flux there. At a domain face the clamped ghosts make every limiter's slope zero, so
// both reconstructed states equal the boundary cell's primitive state and the face flux is face_flux(q, q) with
// that cell's c^2: the value the stage adds to the cell, in the same float arithmetic. Over a stage
// (a = 0, b = 1), sum rho_out = sum rho - (dt / dx) out[1] up to the rounding of each cell's update and the density
// floor, since every interior face flux enters two cells with opposite signs.
__global__ __launch_bounds__(256) void mass_sum_kernel(const float* __restrict__ rho, long long n, double* out) {
  __shared__ double red[4];
  double a = 0.0;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x)
    a += (double)rho[q];
  for (int o = 32; o > 0; o >>= 1) a += __shfl_down(a, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(&out[0], ((red[0] + red[1]) + red[2]) + red[3]);
}

template <int SOLVER>
__device__ __forceinline__ float boundary_mass_flux(const float* rho, const float* mx, const float* my,
                                                    const float* mz, const float* xi, long long c, int axis,
                                                    float floor_, float cn2, float dci) {
  const float r = fmaxf(rho[c], floor_);
  const float m[3] = {mx[c], my[c], mz[c]};
  float v[3];
  div_rn<3>(m, recip(r), v);
  const float4 q = make_float4(r, v[0], v[1], v[2]);
  const float c2 = xi[c] * dci + cn2;
  const float cs = sqrtf(c2);
  const float4 F = axis == 0   ? face_flux<Exact, 0, SOLVER>(q, q, c2, c2, cs, cs, floor_)
                   : axis == 1 ? face_flux<Exact, 1, SOLVER>(q, q, c2, c2, cs, cs, floor_)
                               : face_flux<Exact, 2, SOLVER>(q, q, c2, c2, cs, cs, floor_);
  return F.x;
}

template <int SOLVER>
__global__ __launch_bounds__(256) void boundary_flux_kernel(const float* __restrict__ rho, const float* __restrict__ mx,
                                                            const float* __restrict__ my, const float* __restrict__ mz,
                                                            const float* __restrict__ xi, int nx, int ny, int nz,
                                                            float floor_, float cn2, float dci, double* out) {
  __shared__ double red[4];
  const long long fx = (long long)ny * nz, fy = (long long)nx * nz, fz = (long long)nx * ny;
  const long long total = 2 * (fx + fy + fz);
  double a = 0.0;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    long long u = t;
    int axis, side, i, j, k;
    if (u < 2 * fx) {
      axis = 0; side = (int)(u / fx); u %= fx;
      i = side ? nx - 1 : 0; j = (int)(u / nz); k = (int)(u % nz);
    } else if ((u -= 2 * fx) < 2 * fy) {
      axis = 1; side = (int)(u / fy); u %= fy;
      j = side ? ny - 1 : 0; i = (int)(u / nz); k = (int)(u % nz);
    } else {
      u -= 2 * fy;
      axis = 2; side = (int)(u / fz); u %= fz;
      k = side ? nz - 1 : 0; i = (int)(u / ny); j = (int)(u % ny);
    }
    const long long c = cell(i, j, k, ny, nz);
    const float f = boundary_mass_flux<SOLVER>(rho, mx, my, mz, xi, c, axis, floor_, cn2, dci);
    a += side ? (double)f : -(double)f;       // outward: + on the high faces, - on the low faces
  }
  for (int o = 32; o > 0; o >>= 1) a += __shfl_down(a, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(&out[1], ((red[0] + red[1]) + red[2]) + red[3]);
}

// Device-side timestep of n grids (Sim._compute_dt, sim.py:1009-1012, and the clock of sim.py:1099-1114), in
// the host's double arithmetic: dt = (cfl dx) / smax, capped, rounded to float for the stages; t += dt; and smax
// is cleared for the next reduction.
__global__ void cfl_dt_kernel(float* smax, const double* __restrict__ cfl_dx, const double* __restrict__ dt_cap,
                              float* dt_f, double* dt_d, double* t, int n) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  double dt = cfl_dx[s] / (double)smax[s];
  if (dt_cap && dt_cap[s] > 0.0 && dt_cap[s] < dt) dt = dt_cap[s];
  dt_f[s] = (float)dt;
  if (dt_d) dt_d[s] = dt;
  if (t) t[s] = t[s] + dt;
  smax[s] = 0.0f;
}

// max_speed (nebula3d.py:154-162) of float32 (rho, m) fields
__global__ __launch_bounds__(256) void max_speed_kernel(const float* rho, const float* mx, const float* my,
                                                        const float* mz, const float* xi, long long n, float cn2,
                                                        float ci2, float* out) {
  __shared__ float red[4];
  const float dci = ci2 - cn2;
  float m = 0.0f;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x) {
    const float r = rho[q];
    const float s = ((fabsf(mx[q] / r) + fabsf(my[q] / r)) + fabsf(mz[q] / r)) + sqrtf(cn2 + xi[q] * dci);
    m = fmaxf(m, s);
  }
  block_max_atomic(m, red, out);
}

// the same for binary16 (rho, v) fields: the velocity is stored
__global__ __launch_bounds__(256) void max_speed_h_kernel(const uint16_t* vx, const uint16_t* vy, const uint16_t* vz,
                                                          const float* xi, long long n, float cn2, float ci2,
                                                          float* out) {
  __shared__ float red[4];
  const float dci = ci2 - cn2;
  float m = 0.0f;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x) {
    const float s = ((fabsf(h2f(vx[q])) + fabsf(h2f(vy[q]))) + fabsf(h2f(vz[q]))) + sqrtf(cn2 + xi[q] * dci);
    m = fmaxf(m, s);
  }
  block_max_atomic(m, red, out);
}

using HydroKernel = void (*)(HydroArgs);

template <class TL, class ST, class AR, bool SMAX, bool STAGE1>
static HydroKernel pick_lim(int solver, int lim) {
  if (solver == NV_SOLVER_RUSANOV) return hydro_stage_kernel<TL, ST, AR, NV_SOLVER_RUSANOV, NV_LIM_MINMOD, SMAX, STAGE1>;
  if (lim == NV_LIM_VANLEER) return hydro_stage_kernel<TL, ST, AR, NV_SOLVER_HLL, NV_LIM_VANLEER, SMAX, STAGE1>;
  if (lim == NV_LIM_MC) return hydro_stage_kernel<TL, ST, AR, NV_SOLVER_HLL, NV_LIM_MC, SMAX, STAGE1>;
  return hydro_stage_kernel<TL, ST, AR, NV_SOLVER_HLL, NV_LIM_MINMOD, SMAX, STAGE1>;
}

template <class TL, class ST, class AR>
static HydroKernel pick(int solver, int lim, bool smax, bool stage1) {
  if (smax) return stage1 ? pick_lim<TL, ST, AR, true, true>(solver, lim) : pick_lim<TL, ST, AR, true, false>(solver, lim);
  return stage1 ? pick_lim<TL, ST, AR, false, true>(solver, lim) : pick_lim<TL, ST, AR, false, false>(solver, lim);
}

enum class Mode { F32_EXACT, F32_FAST, F16V_FAST };

template <class TL>
static HydroKernel pick_mode_tl(Mode mode, int solver, int lim, bool sm, bool stage1) {
  return mode == Mode::F32_EXACT  ? pick<TL, F32, Exact>(solver, lim, sm, stage1)
         : mode == Mode::F32_FAST ? pick<TL, F32, Fast>(solver, lim, sm, stage1)
                                  : pick<TL, F16V, Fast>(solver, lim, sm, stage1);
}

static HydroKernel pick_mode(int tile, Mode mode, int solver, int lim, bool sm, bool stage1) {
  return tile == TILE_32x16   ? pick_mode_tl<T32>(mode, solver, lim, sm, stage1)
         : tile == TILE_64x16 ? pick_mode_tl<T6416>(mode, solver, lim, sm, stage1)
                              : pick_mode_tl<T64>(mode, solver, lim, sm, stage1);
}

static int tile_tk(int tile) { return tile == TILE_32x16 ? T32::TK : tile == TILE_64x16 ? T6416::TK : T64::TK; }
static int tile_tj(int tile) { return tile == TILE_32x16 ? T32::TJ : tile == TILE_64x16 ? T6416::TJ : T64::TJ; }

// The shape when the caller leaves it to the library, by idle lanes: 64 x 8 when nz is a multiple of 64, 32 x 16
// when it is a multiple of 32 only, otherwise the one with fewer idle lanes. 64 x 16 (one 1024-thread workgroup per
// CU) is not chosen automatically: at 512^3 it is slower in every form (2.32 / 2.70 ms against 2.21 / 2.59 for the
// exact stages).
static int auto_tile(int nz) {
  if (nz % 64 == 0) return TILE_64x8;
  const int idle64 = ((nz + 63) / 64) * 64 - nz, idle32 = ((nz + 31) / 32) * 32 - nz;
  return idle32 < idle64 ? TILE_32x16 : TILE_64x8;
}

// Validates one grid's stage arguments and fills its kernel arguments. jk_total: the number of (y, z) tile
// columns the launch covers (this grid's own, or a whole batch's), which sets the march depth ci so that the
// launch has about 2048 x-chunks of tiles in all, at most max_ci planes each.
static int fill_stage(const char* name, const NvStageDesc& d, int solver, int lim, int tile, long long jk_total,
                      HydroArgs& A, int depth, int max_ci) {
  NV_REQUIRE(d.u[0] && d.u[1] && d.u[2] && d.u[3] && d.xi && d.u0[0] && d.u0[1] && d.u0[2] && d.u0[3] && d.out[0] &&
                 d.out[1] && d.out[2] && d.out[3],
             "%s: null field pointer", name);
  NV_REQUIRE(d.nx > 0 && d.ny > 0 && d.nz > 0, "%s: bad grid %dx%dx%d", name, d.nx, d.ny, d.nz);
  NV_REQUIRE(d.nstars >= 0 && (d.nstars == 0 || (d.stars && d.jets)), "%s: nstars=%d needs star/jet tables", name,
             d.nstars);
  NV_REQUIRE(solver == NV_SOLVER_RUSANOV || solver == NV_SOLVER_HLL, "%s: bad solver %d", name, solver);
  NV_REQUIRE(lim >= NV_LIM_MINMOD && lim <= NV_LIM_MC, "%s: bad limiter %d", name, lim);
  NV_REQUIRE(d.out[0] != d.u[0] && d.out[1] != d.u[1] && d.out[2] != d.u[2] && d.out[3] != d.u[3],
             "%s: output aliases the input U", name);
  NV_REQUIRE(d.floor_ > 0.0f && d.dx > 0.0f, "%s: floor and dx must be > 0", name);
  NV_REQUIRE((long long)d.ny * d.nz < (1LL << 30), "%s: ny nz = %lld, at or above 2^30 cells per x plane", name,
             (long long)d.ny * d.nz);
  A.f0 = d.u[0]; A.f1 = d.u[1]; A.f2 = d.u[2]; A.f3 = d.u[3]; A.xi = d.xi;
  A.g0 = d.u0[0]; A.g1 = d.u0[1]; A.g2 = d.u0[2]; A.g3 = d.u0[3];
  A.o0 = d.out[0]; A.o1 = d.out[1]; A.o2 = d.out[2]; A.o3 = d.out[3];
  A.stars = reinterpret_cast<const float4*>(d.stars);
  A.jets = reinterpret_cast<const float4*>(d.jets);
  A.smax = d.smax_out;
  A.dt_dev = d.dt_dev;
  A.nx = d.nx; A.ny = d.ny; A.nz = d.nz; A.nstars = d.nstars;
  A.u0_is_u = (d.u0[0] == d.u[0] && d.u0[1] == d.u[1] && d.u0[2] == d.u[2] && d.u0[3] == d.u[3]) ? 1 : 0;
  A.wind = d.wind; A.jet_cos = d.jet_cos; A.jet_r = d.jet_r; A.pulse = d.pulse; A.dt = d.dt; A.dx = d.dx;
  A.a = d.a; A.b = d.b; A.floor_ = d.floor_; A.cn2 = d.cn2; A.ci2 = d.ci2; A.vmax = d.vmax;
  {
    const double v2 = (double)d.vmax * (double)d.vmax;
    float lo = (float)v2;
    if ((double)lo > v2) lo = nextafterf(lo, 0.0f);
    A.vmax2_lo = lo;
  }
  A.inv_dx = pow2_reciprocal(d.dx);
  {
    // jet_r^2 (1 + 2^-20), rounded up: exceeds the real jet_r^2 by far more than the rounding of either
    const double j2 = (double)d.jet_r * (double)d.jet_r * (1.0 + 0x1p-20);
    float h = (float)j2;
    if ((double)h < j2) h = nextafterf(h, INFINITY);
    A.jet_r2_hi = d.jet_r > 0.0f ? h : -1.0f;
  }
  // march depth: the caller's, or long enough to amortise the 4-plane prologue and short enough to fill the chip
  A.ci = depth > 0 ? std::min(depth, d.nx)
                   : (int)std::min<long long>(max_ci, std::max<long long>(4, (long long)d.nx * jk_total / 2048));
  A.tiles_k = (d.nz + tile_tk(tile) - 1) / tile_tk(tile);
  A.tiles_j = (d.ny + tile_tj(tile) - 1) / tile_tj(tile);
  A.tiles = A.tiles_k * A.tiles_j * ((d.nx + A.ci - 1) / A.ci);
  A.per_xcd = (A.tiles + 7) / 8;
  A.run = A.per_xcd;
  return NV_OK;
}

static long long tiles_jk(int tile, int ny, int nz) {
  return (long long)((nz + tile_tk(tile) - 1) / tile_tk(tile)) * ((ny + tile_tj(tile) - 1) / tile_tj(tile));
}

static bool is_stage1(const HydroArgs& A) { return A.u0_is_u && A.a == 0.0f && A.b == 1.0f; }

// workgroups of kern (nt threads) one XCD holds at once: its CUs times the kernel's occupancy per CU; 0 when the
// device cannot be queried (the launch then keeps one contiguous run per XCD). Cached per kernel and device.
static int resident_per_xcd(HydroKernel kern, int nt) {
  static std::mutex mu;
  static std::map<std::pair<const void*, int>, int> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  std::lock_guard<std::mutex> lock(mu);
  const auto key = std::make_pair(reinterpret_cast<const void*>(kern), dev);
  const auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  int cus = 0, per_cu = 0;
  int r = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, nt, 0) == hipSuccess && per_cu > 0)
    r = cus / 8 * per_cu;
  cache[key] = r;
  return r;
}

static int launch_desc(Mode mode, const char* name, const NvStageDesc& d, int solver, int lim, int tile,
                       void* stream) {
  const int depth = tile >> 8;   // x planes per workgroup (NV_TILE_DEPTH), 0 for the library's choice
  tile &= 0xff;
  NV_REQUIRE(tile >= TILE_AUTO && tile <= TILE_64x16 && depth >= 0, "%s: bad tile shape %d (depth %d)", name, tile,
             depth);
  if (tile == TILE_AUTO) tile = auto_tile(d.nz);
  HydroArgs A;
  const int rc = fill_stage(name, d, solver, lim, tile, tiles_jk(tile, d.ny, d.nz), A, depth, 64);
  if (rc != NV_OK) return rc;
  HydroKernel kern = pick_mode(tile, mode, solver, lim, d.smax_out != nullptr, is_stage1(A));
  // more tiles than one round of resident workgroups: bands of as many tiles as an XCD holds (band_tile); the
  // launch then covers whole bands (the tiles past A.tiles return at once)
  const int per_xcd_res = resident_per_xcd(kern, tile_tk(tile) * tile_tj(tile));
  if (per_xcd_res > 0 && A.per_xcd > per_xcd_res) {
    A.run = per_xcd_res;
    A.per_xcd = (A.per_xcd + per_xcd_res - 1) / per_xcd_res * per_xcd_res;
  }
  hipLaunchKernelGGL(kern, dim3(8 * A.per_xcd), dim3(tile_tk(tile), tile_tj(tile), 1), 0, as_stream(stream), A);
  return check_launch(name);
}

static int launch_stage(Mode mode, const char* name, const void* f0, const void* f1, const void* f2, const void* f3,
                        const float* xi, const void* g0, const void* g1, const void* g2, const void* g3, void* o0,
                        void* o1, void* o2, void* o3, int nx, int ny, int nz, const float* stars, int nstars,
                        float wind, const float* jets, float jet_cos, float jet_r, float pulse, float dt, float dx,
                        float a, float b, float floor_, float cn2, float ci2, float vmax, int solver, int lim,
                        float* smax_out, void* stream) {
  clear_error();
  NvStageDesc d = {};
  d.u[0] = f0; d.u[1] = f1; d.u[2] = f2; d.u[3] = f3; d.xi = xi;
  d.u0[0] = g0; d.u0[1] = g1; d.u0[2] = g2; d.u0[3] = g3;
  d.out[0] = o0; d.out[1] = o1; d.out[2] = o2; d.out[3] = o3;
  d.stars = stars; d.jets = jets; d.smax_out = smax_out; d.dt_dev = nullptr;
  d.nx = nx; d.ny = ny; d.nz = nz; d.nstars = nstars;
  d.wind = wind; d.jet_cos = jet_cos; d.jet_r = jet_r; d.pulse = pulse; d.dt = dt; d.dx = dx;
  d.a = a; d.b = b; d.floor_ = floor_; d.cn2 = cn2; d.ci2 = ci2; d.vmax = vmax;
  return launch_desc(mode, name, d, solver, lim, TILE_AUTO, stream);
}

// ------------------------------------------------------------------------------------------ batch plans
// A plan is one host-built blob, uploaded once and used by every launch: header, kernel arguments of each
// member, then the tile prefix start[0..n] (start[n] = total tiles).
// The tile queue's nine counters (zero in the host blob; the persistent launch leaves them zero) follow start[].
struct BatchHeader {
  uint32_t magic;
  int n, mode, solver, lim, smax, stage1, total, per_xcd, tile;
  int pad[6];
};
static_assert(sizeof(BatchHeader) == 64, "plan header layout");
constexpr uint32_t kBatchMagic = 0x4e564231u;   // "NVB1"
constexpr int kQueueInts = 16;                   // nine counters, padded

static long long plan_bytes(int n) {
  return (long long)sizeof(BatchHeader) + (long long)n * (long long)sizeof(HydroArgs) + 4LL * (n + 1) +
         4LL * kQueueInts;
}

typedef void (*BatchKernel)(const HydroArgs*, const int*, int, int, int);
typedef void (*QueueKernel)(const HydroArgs*, const int*, int, int, int, int*);

// the persistent form is built for HLL + minmod (the benchmark workloads' scheme), both stage forms
template <class TL, class ST, class AR>
static QueueKernel pick_queue_st(bool sm, bool stage1) {
  if (sm) return stage1 ? hydro_stage_queue_kernel<TL, ST, AR, NV_SOLVER_HLL, NV_LIM_MINMOD, true, true>
                        : hydro_stage_queue_kernel<TL, ST, AR, NV_SOLVER_HLL, NV_LIM_MINMOD, true, false>;
  return stage1 ? hydro_stage_queue_kernel<TL, ST, AR, NV_SOLVER_HLL, NV_LIM_MINMOD, false, true>
                : hydro_stage_queue_kernel<TL, ST, AR, NV_SOLVER_HLL, NV_LIM_MINMOD, false, false>;
}

template <class TL>
static QueueKernel pick_queue_tl(int mode, bool sm, bool stage1) {
  return mode == NV_PREC_F32        ? pick_queue_st<TL, F32, Exact>(sm, stage1)
         : mode == NV_PREC_F32_FAST ? pick_queue_st<TL, F32, Fast>(sm, stage1)
                                    : pick_queue_st<TL, F16V, Fast>(sm, stage1);
}

template <class TL, class ST, class AR, bool SMAX, bool STAGE1>
static BatchKernel pick_batch_lim(int solver, int lim) {
  if (solver == NV_SOLVER_RUSANOV)
    return hydro_stage_batch_kernel<TL, ST, AR, NV_SOLVER_RUSANOV, NV_LIM_MINMOD, SMAX, STAGE1>;
  if (lim == NV_LIM_VANLEER) return hydro_stage_batch_kernel<TL, ST, AR, NV_SOLVER_HLL, NV_LIM_VANLEER, SMAX, STAGE1>;
  if (lim == NV_LIM_MC) return hydro_stage_batch_kernel<TL, ST, AR, NV_SOLVER_HLL, NV_LIM_MC, SMAX, STAGE1>;
  return hydro_stage_batch_kernel<TL, ST, AR, NV_SOLVER_HLL, NV_LIM_MINMOD, SMAX, STAGE1>;
}

template <class TL, class ST, class AR>
static BatchKernel pick_batch_st(int solver, int lim, bool sm, bool stage1) {
  if (sm)
    return stage1 ? pick_batch_lim<TL, ST, AR, true, true>(solver, lim)
                  : pick_batch_lim<TL, ST, AR, true, false>(solver, lim);
  return stage1 ? pick_batch_lim<TL, ST, AR, false, true>(solver, lim)
                : pick_batch_lim<TL, ST, AR, false, false>(solver, lim);
}

template <class TL>
static BatchKernel pick_batch_tl(int mode, int solver, int lim, bool sm, bool stage1) {
  return mode == NV_PREC_F32        ? pick_batch_st<TL, F32, Exact>(solver, lim, sm, stage1)
         : mode == NV_PREC_F32_FAST ? pick_batch_st<TL, F32, Fast>(solver, lim, sm, stage1)
                                    : pick_batch_st<TL, F16V, Fast>(solver, lim, sm, stage1);
}

// The shape for a batch: the one with fewer idle lanes over all members (cells of the padded z extents).
static int auto_tile_batch(const NvStageDesc* m, int n) {
  long long pad64 = 0, pad32 = 0;
  for (int s = 0; s < n; ++s) {
    const long long xy = (long long)m[s].nx * m[s].ny;
    pad64 += xy * (((m[s].nz + 63) / 64) * 64);
    pad32 += xy * (((m[s].nz + 31) / 32) * 32);
  }
  return pad32 < pad64 ? TILE_32x16 : TILE_64x8;
}

}  // namespace nv

#define NV_STAGE_TAIL                                                                                      \
  nx, ny, nz, stars, nstars, wind, jets, jet_cos, jet_r, pulse, dt, dx, a, b, floor_, cn2, ci2, vmax, solver, lim, \
      smax_out, stream

extern "C" int nv_hydro_stage(const float* rho, const float* mx, const float* my, const float* mz, const float* xi,
                              const float* r0, const float* mx0, const float* my0, const float* mz0, float* o_r,
                              float* o_mx, float* o_my, float* o_mz, int nx, int ny, int nz, const float* stars,
                              int nstars, float wind, const float* jets, float jet_cos, float jet_r, float pulse,
                              float dt, float dx, float a, float b, float floor_, float cn2, float ci2, float vmax,
                              int solver, int lim, float* smax_out, void* stream) {
  return nv::launch_stage(nv::Mode::F32_EXACT, "nv_hydro_stage", rho, mx, my, mz, xi, r0, mx0, my0, mz0, o_r, o_mx,
                          o_my, o_mz, NV_STAGE_TAIL);
}

extern "C" int nv_hydro_stage_fast(const float* rho, const float* mx, const float* my, const float* mz,
                                   const float* xi, const float* r0, const float* mx0, const float* my0,
                                   const float* mz0, float* o_r, float* o_mx, float* o_my, float* o_mz, int nx, int ny,
                                   int nz, const float* stars, int nstars, float wind, const float* jets,
                                   float jet_cos, float jet_r, float pulse, float dt, float dx, float a, float b,
                                   float floor_, float cn2, float ci2, float vmax, int solver, int lim,
                                   float* smax_out, void* stream) {
  return nv::launch_stage(nv::Mode::F32_FAST, "nv_hydro_stage_fast", rho, mx, my, mz, xi, r0, mx0, my0, mz0, o_r,
                          o_mx, o_my, o_mz, NV_STAGE_TAIL);
}

extern "C" int nv_hydro_stage_h(const float* rho, const uint16_t* vx, const uint16_t* vy, const uint16_t* vz,
                                const float* xi, const float* r0, const uint16_t* vx0, const uint16_t* vy0,
                                const uint16_t* vz0, float* o_r, uint16_t* o_vx, uint16_t* o_vy, uint16_t* o_vz,
                                int nx, int ny, int nz, const float* stars, int nstars, float wind, const float* jets,
                                float jet_cos, float jet_r, float pulse, float dt, float dx, float a, float b,
                                float floor_, float cn2, float ci2, float vmax, int solver, int lim, float* smax_out,
                                void* stream) {
  return nv::launch_stage(nv::Mode::F16V_FAST, "nv_hydro_stage_h", rho, vx, vy, vz, xi, r0, vx0, vy0, vz0, o_r, o_vx,
                          o_vy, o_vz, NV_STAGE_TAIL);
}

extern "C" int nv_hydro_stage_ex(const NvStageDesc* d, int precision, int solver, int lim, int tile, void* stream) {
  using namespace nv;
  clear_error();
  NV_REQUIRE(d, "nv_hydro_stage_ex: null descriptor");
  NV_REQUIRE(precision >= NV_PREC_F32 && precision <= NV_PREC_F16V, "nv_hydro_stage_ex: bad precision %d", precision);
  const Mode mode = precision == NV_PREC_F32 ? Mode::F32_EXACT : precision == NV_PREC_F32_FAST ? Mode::F32_FAST
                                                                                               : Mode::F16V_FAST;
  return launch_desc(mode, "nv_hydro_stage_ex", *d, solver, lim, tile, stream);
}

extern "C" long long nv_hydro_stage_batch_plan_size(int n) { return n > 0 ? nv::plan_bytes(n) : -1; }

extern "C" int nv_hydro_stage_batch_plan(const NvStageDesc* members, int n, int precision, int solver, int lim,
                                         int tile, void* plan, long long plan_bytes) {
  using namespace nv;
  clear_error();
  NV_REQUIRE(members && plan && n > 0, "nv_hydro_stage_batch_plan: need n > 0 members and a plan buffer");
  NV_REQUIRE(plan_bytes >= nv::plan_bytes(n), "nv_hydro_stage_batch_plan: plan needs %lld bytes, got %lld",
             nv::plan_bytes(n), plan_bytes);
  NV_REQUIRE(precision >= NV_PREC_F32 && precision <= NV_PREC_F16V, "nv_hydro_stage_batch_plan: bad precision %d",
             precision);
  const int depth = tile >> 8;   // x planes per workgroup (NV_TILE_DEPTH; each member's nx at most)
  tile &= 0xff;
  NV_REQUIRE(tile >= TILE_AUTO && tile <= TILE_64x16 && depth >= 0,
             "nv_hydro_stage_batch_plan: bad tile shape %d (depth %d)", tile, depth);
  long long jk_total = 0;
  for (int s = 0; s < n; ++s) {
    NV_REQUIRE(members[s].nx > 0 && members[s].ny > 0 && members[s].nz > 0,
               "nv_hydro_stage_batch_plan: member %d: bad grid", s);
  }
  if (tile == TILE_AUTO) tile = auto_tile_batch(members, n);
  for (int s = 0; s < n; ++s) jk_total += tiles_jk(tile, members[s].ny, members[s].nz);
  BatchHeader H = {};
  HydroArgs* args = reinterpret_cast<HydroArgs*>(static_cast<char*>(plan) + sizeof(BatchHeader));
  int* start = reinterpret_cast<int*>(args + n);
  const bool sm = members[0].smax_out != nullptr;
  bool stage1 = true;
  long long total = 0;
  char name[64];
  for (int s = 0; s < n; ++s) {
    snprintf(name, sizeof(name), "nv_hydro_stage_batch_plan: member %d", s);
    HydroArgs A;
    // the library's depth for a batch: at most 32 planes (the 64-member batch's fast stages 0.93 / 1.06 ms per
    // launch against 0.96 / 1.10 at up to 64, 0.95 / 1.08 at 16, 1.00 / 1.16 at 8; the same bits)
    const int rc = fill_stage(name, members[s], solver, lim, tile, jk_total, A, depth, 32);
    if (rc != NV_OK) return rc;
    NV_REQUIRE((members[s].smax_out != nullptr) == sm, "%s: smax_out must be set for all members or none", name);
    stage1 = stage1 && is_stage1(A);
    start[s] = (int)total;
    total += A.tiles;
    NV_REQUIRE(total < (1LL << 30), "nv_hydro_stage_batch_plan: more than 2^30 tiles");
    memcpy(&args[s], &A, sizeof(A));
  }
  start[n] = (int)total;
  memset(start + n + 1, 0, 4 * kQueueInts);
  H.magic = kBatchMagic;
  H.n = n;
  H.mode = precision;
  H.solver = solver;
  H.lim = lim;
  H.smax = sm ? 1 : 0;
  H.stage1 = stage1 ? 1 : 0;
  H.total = (int)total;
  H.per_xcd = (int)((total + 7) / 8);
  H.tile = tile;
  memcpy(plan, &H, sizeof(H));
  return NV_OK;
}

namespace nv {
// Hardware-dispatched (one workgroup per tile) or persistent (a queue-fed grid of resident workgroups) launch.
// wgs_per_cu: resident workgroups per CU of the persistent grid (its waves per CU divided by the tile's wavefronts),
// at most what the kernel's registers and LDS allow; 0 for that maximum.
static int batch_launch(const char* name, const void* plan_host, const void* plan_dev, bool persistent,
                        int wgs_per_cu, void* stream) {
  clear_error();
  NV_REQUIRE(plan_host && plan_dev, "%s: null plan", name);
  NV_REQUIRE(wgs_per_cu >= 0, "%s: wgs_per_cu = %d must be >= 0", name, wgs_per_cu);
  BatchHeader H;
  memcpy(&H, plan_host, sizeof(H));
  NV_REQUIRE(H.magic == kBatchMagic && H.n > 0, "%s: not a stage batch plan", name);
  const HydroArgs* args = reinterpret_cast<const HydroArgs*>(static_cast<const char*>(plan_dev) + sizeof(BatchHeader));
  const int* start = reinterpret_cast<const int*>(args + H.n);
  const bool sm = H.smax != 0, st1 = H.stage1 != 0;
  const dim3 block(tile_tk(H.tile), tile_tj(H.tile), 1);
  if (!persistent) {
    BatchKernel kern = H.tile == TILE_32x16   ? pick_batch_tl<T32>(H.mode, H.solver, H.lim, sm, st1)
                       : H.tile == TILE_64x16 ? pick_batch_tl<T6416>(H.mode, H.solver, H.lim, sm, st1)
                                              : pick_batch_tl<T64>(H.mode, H.solver, H.lim, sm, st1);
    hipLaunchKernelGGL(kern, dim3(8 * H.per_xcd), block, 0, as_stream(stream), args, start, H.n, H.total, H.per_xcd);
    return check_launch(name);
  }
  NV_REQUIRE(H.solver == NV_SOLVER_HLL && H.lim == NV_LIM_MINMOD, "%s: the persistent form is built for HLL + minmod",
             name);
  int* queue = const_cast<int*>(start + H.n + 1);
  QueueKernel kern = H.tile == TILE_32x16   ? pick_queue_tl<T32>(H.mode, sm, st1)
                     : H.tile == TILE_64x16 ? pick_queue_tl<T6416>(H.mode, sm, st1)
                                            : pick_queue_tl<T64>(H.mode, sm, st1);
  // as many workgroups as the device holds at once (a multiple of 8, one share per XCD), never more than tiles
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
      hipSuccess || hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, (int)(block.x * block.y), 0) !=
      hipSuccess) {
    set_error("%s: cannot query the device's occupancy", name);
    return NV_ERR_LAUNCH;
  }
  const int fit = std::max(per_cu, 1);
  const int use = wgs_per_cu > 0 ? std::min(wgs_per_cu, fit) : fit;
  const long long resident = std::max<long long>(8, ((long long)cus * use) / 8 * 8);
  const long long grid = std::min<long long>(resident, ((long long)H.total + 7) / 8 * 8);
  hipLaunchKernelGGL(kern, dim3((unsigned)grid), block, 0, as_stream(stream), args, start, H.n, H.total, H.per_xcd,
                     queue);
  return check_launch(name);
}
}  // namespace nv

extern "C" int nv_hydro_stage_batch(const void* plan_host, const void* plan_dev, void* stream) {
  return nv::batch_launch("nv_hydro_stage_batch", plan_host, plan_dev, false, 0, stream);
}

extern "C" int nv_hydro_stage_batch_queue(const void* plan_host, const void* plan_dev, void* stream) {
  return nv::batch_launch("nv_hydro_stage_batch_queue", plan_host, plan_dev, true, 0, stream);
}

extern "C" int nv_hydro_stage_batch_queue_ex(const void* plan_host, const void* plan_dev, int wgs_per_cu,
                                             void* stream) {
  return nv::batch_launch("nv_hydro_stage_batch_queue_ex", plan_host, plan_dev, true, wgs_per_cu, stream);
}

extern "C" int nv_conservation_sums(const float* rho, const float* mx, const float* my, const float* mz,
                                    const float* xi, int nx, int ny, int nz, float floor_, float cn2, float ci2,
                                    int solver, double* out, void* stream) {
  using namespace nv;
  clear_error();
  NV_REQUIRE(rho && mx && my && mz && xi && out, "nv_conservation_sums: null pointer");
  NV_REQUIRE(nx > 0 && ny > 0 && nz > 0, "nv_conservation_sums: bad grid");
  NV_REQUIRE(solver == NV_SOLVER_RUSANOV || solver == NV_SOLVER_HLL, "nv_conservation_sums: bad solver %d", solver);
  NV_REQUIRE(floor_ > 0.0f, "nv_conservation_sums: floor must be > 0");
  const long long n = (long long)nx * ny * nz;
  hipLaunchKernelGGL(mass_sum_kernel, dim3((int)std::min<long long>((n + 255) / 256, 4096)), dim3(256), 0,
                     as_stream(stream), rho, n, out);
  const long long faces = 2 * ((long long)ny * nz + (long long)nx * nz + (long long)nx * ny);
  const int blocks = (int)std::min<long long>((faces + 255) / 256, 4096);
  auto kern = solver == NV_SOLVER_HLL ? boundary_flux_kernel<NV_SOLVER_HLL> : boundary_flux_kernel<NV_SOLVER_RUSANOV>;
  hipLaunchKernelGGL(kern, dim3(blocks), dim3(256), 0, as_stream(stream), rho, mx, my, mz, xi, nx, ny, nz, floor_, cn2,
                     ci2 - cn2, out);
  return check_launch("nv_conservation_sums");
}

extern "C" int nv_cfl_dt(float* smax, const double* cfl_dx, const double* dt_cap, float* dt_f, double* dt_d,
                         double* t, int n, void* stream) {
  using namespace nv;
  clear_error();
  NV_REQUIRE(smax && cfl_dx && dt_f && n > 0, "nv_cfl_dt: null pointer or n <= 0");
  hipLaunchKernelGGL(cfl_dt_kernel, dim3((n + 255) / 256), dim3(256), 0, as_stream(stream), smax, cfl_dx, dt_cap, dt_f,
                     dt_d, t, n);
  return check_launch("nv_cfl_dt");
}

extern "C" int nv_max_speed(const float* rho, const float* mx, const float* my, const float* mz, const float* xi,
                            int nx, int ny, int nz, float cn2, float ci2, float* out, void* stream) {
  using namespace nv;
  clear_error();
  NV_REQUIRE(rho && mx && my && mz && xi && out, "nv_max_speed: null pointer");
  NV_REQUIRE(nx > 0 && ny > 0 && nz > 0, "nv_max_speed: bad grid %dx%dx%d", nx, ny, nz);
  const long long n = (long long)nx * ny * nz;
  const int blocks = (int)std::min<long long>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(max_speed_kernel, dim3(blocks), dim3(256), 0, as_stream(stream), rho, mx, my, mz, xi, n, cn2,
                     ci2, out);
  return check_launch("nv_max_speed");
}

extern "C" int nv_max_speed_h(const uint16_t* vx, const uint16_t* vy, const uint16_t* vz, const float* xi, int nx,
                              int ny, int nz, float cn2, float ci2, float* out, void* stream) {
  using namespace nv;
  clear_error();
  NV_REQUIRE(vx && vy && vz && xi && out, "nv_max_speed_h: null pointer");
  NV_REQUIRE(nx > 0 && ny > 0 && nz > 0, "nv_max_speed_h: bad grid %dx%dx%d", nx, ny, nz);
  const long long n = (long long)nx * ny * nz;
  const int blocks = (int)std::min<long long>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(max_speed_h_kernel, dim3(blocks), dim3(256), 0, as_stream(stream), vx, vy, vz, xi, n, cn2, ci2,
                     out);
  return check_launch("nv_max_speed_h");
}
