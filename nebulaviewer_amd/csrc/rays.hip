// rays.hip - Stromgren ionization and FUV attenuation for gfx950.
//
// Reference: sample_n (nebula3d.py:195-226), box_range (229-244), cast_rays (247-270), rf_at (273-278),
// apply_front (281-307), smooth_xi (310-315), cast_fuv (320-352), apply_fuv (355-388); the per-call
// sequences of Sim.ionize (sim.py:927-947) and Sim.fuv_field (sim.py:949-978); to_half (sim.py:511-514).
//
// The casts are gather-bound (8 trilinear corner loads per sample, one thread per ray, 64 neighbouring
// directions per wavefront so their samples share cache lines). The per-cell passes are fused:
// nv_ionize_apply does zero/fill + every star's front lookup + the 7-point smoothing + the fp16 store in one
// pass (a 1-cell halo of the unsmoothed xi is recomputed in LDS instead of being written and re-read), and
// nv_fuv_apply accumulates every star into a register before the single store.
// Transcendentals come from nv_math.h so the CPU oracle reproduces them bit for bit.
#include <string.h>
#include <algorithm>
#include <cmath>
#include <mutex>
#include <type_traits>
#include "nv_common.h"
#include "nv_math.h"
#include "nv_divide.h"
#include "nv_sample.h"

namespace nv {

constexpr float PI_F = 3.14159265358979323846f;

// slab test against [0, bmax] (nebula3d.py:229-244)
__device__ __forceinline__ void box_range(const float o[3], const float d[3], const float bm[3], float* t0o,
                                          float* t1o) {
  float t0 = 0.0f, t1 = 1e9f;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float da = d[a];
    if (fabsf(da) < 1e-9f) {
      if (o[a] < 0.0f || o[a] > bm[a]) t1 = -1.0f;
    } else {
      const float ta = (0.0f - o[a]) / da, tb = (bm[a] - o[a]) / da;
      t0 = fmaxf(t0, fminf(ta, tb));
      t1 = fminf(t1, fmaxf(ta, tb));
    }
  }
  *t0o = t0;
  *t1o = t1;
}

// box_range for a star strictly inside the box: each axis's entry quotient is negative, so t0 = 0, and t1 is the
// least of the exit quotients, the same divisions box_range takes its maximum of (three divisions instead of six)
__device__ __forceinline__ float box_exit_inside(const float o[3], const float d[3], const float bm[3]) {
  float t1 = 1e9f;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float da = d[a];
    if (fabsf(da) >= 1e-9f) t1 = fminf(t1, da > 0.0f ? (bm[a] - o[a]) / da : (0.0f - o[a]) / da);
  }
  return t1;
}

// direction of angular cell (iu, ip) (nebula3d.py:252-257)
__device__ __forceinline__ void ray_dir(int iu, int ip, int nu, int nphi, float d[3]) {
  const float u = -1.0f + ((float)iu + 0.5f) * 2.0f / (float)nu;
  const float s = sqrtf(fmaxf(1.0f - u * u, 0.0f));
  const float phi = ((float)ip + 0.5f) * 2.0f * PI_F / (float)nphi;
  float sn, cs;
  nvm::sincosf_det(phi, &sn, &cs);
  d[0] = s * cs;
  d[1] = s * sn;
  d[2] = u;
}

__device__ __forceinline__ float ldf(const float* __restrict__ f, unsigned i) { return f[i]; }
__device__ __forceinline__ float ldf(const uint16_t* __restrict__ f, unsigned i) {
  return (float)__builtin_bit_cast(_Float16, f[i]);
}

// Direction cell of lane `lane` of direction patch `patch`: each wavefront takes an 8 x 8 patch of (u, phi) cells,
// so its 64 rays stay within a few cells of each other and their corner gathers share cache lines. Each quad of
// lanes (4q .. 4q+3), which the L1 serves together, takes a 2 x 2 block of (u, phi) cells rather than a row of
// four: its rays lie closer together and touch fewer cache lines (the casts at 512^3: 1.35 -> 1.22 ms). The
// patch index is the same for the whole wavefront (wave_patch), so its division is scalar arithmetic.
__device__ __forceinline__ bool ray_of_patch(int patch, int lane, int nu, int nphi, int* iu, int* ip) {
  const int tp = (nphi + 7) >> 3;
  const int pu = patch / tp;
  *iu = pu * 8 + ((lane >> 4) << 1) + ((lane >> 1) & 1);
  *ip = (patch - pu * tp) * 8 + (((lane >> 2) & 3) << 1) + (lane & 1);
  return *iu < nu && *ip < nphi;
}
// patch of this wavefront when block b of 256 threads holds patches 4b .. 4b+3
__device__ __forceinline__ int wave_patch(int b) {
  return b * 4 + (int)__builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
}

// One ray-cast job: a star, its direction grid and the density grid it marches through.
struct RayJob {
  const float* rho;
  const float4* bricks;   // nv_ray_bricks of rho (the two-load sampler), or null (the paired sampler)
  const float4* dirs;     // nv_ray_dirs of (nu, nphi): each ray's direction, or null (computed per ray)
  float* rf_range;        // (least, greatest) entry of rfront, written after the casts, or null
  float* rfront;
  Recip rdx;         // dx and 1/dx correctly rounded (the shared reciprocal of the general cell scaling)
  int nx, ny, nz, nsteps, nu, nphi, wide;
  int inside;        // the star lies strictly inside the box (0 < s < b on every axis)
  float sx, sy, sz, Q, dr, dx, rcap, bx, by, bz;
  float inv;         // 1/dx when dx is a power of two, else 0
};

// cast_rays (nebula3d.py:247-270) for the ray of thread t of a job. Two samples per iteration: the corner loads
// of sample n+1 are in flight while sample n is tested, and the column is still accumulated sample by sample in
// the reference order. WIDE: a ray whose marched length (n1 - n0) dr exceeds 512 cells keeps its column in FP64
// (the test col + inc > Q and the interpolation of rf in double); shorter rays are unchanged.
// Returns the ray's front radius (>= +0), or NaN for a lane without a ray.
template <bool PAIR, bool WIDE, class OFF, class SC, bool BRICK>
__device__ __forceinline__ float cast_one(const RayJob& J, const SC& cs, int patch, int lane) {
  int iu, ip;
  if (!ray_of_patch(patch, lane, J.nu, J.nphi, &iu, &ip)) return __builtin_nanf("");
  const int nx = J.nx, ny = J.ny, nz = J.nz;
  const float sx = J.sx, sy = J.sy, sz = J.sz, Q = J.Q, dr = J.dr;
  float d[3];
  if (J.dirs) {
    const float4 dv = J.dirs[iu * J.nphi + ip];
    d[0] = dv.x;
    d[1] = dv.y;
    d[2] = dv.z;
  } else {
    ray_dir(iu, ip, J.nu, J.nphi, d);
  }
  const float st[3] = {sx, sy, sz}, bm[3] = {J.bx, J.by, J.bz};
  float t0, t1;
  int n0 = 0;
  if (J.inside) {
    t1 = box_exit_inside(st, d, bm);   // t0 = 0 and n0 = (int)(0 / dr) = 0
  } else {
    box_range(st, d, bm, &t0, &t1);
    n0 = (int)(t0 / dr);
  }
  const int n1 = min(J.nsteps, (int)(t1 / dr) + 2);
  const bool wide = WIDE && (float)(n1 - n0) * dr > 512.0f * J.dx;
  float col = 0.0f, rf = J.rcap;
  double col64 = 0.0;
  // one sample's threshold test; true when the front is found (rf set)
  auto step = [&](float r, float inc) -> bool {
    if (WIDE && wide) {
      if (col64 + (double)inc > (double)Q) {
        rf = (float)(((double)r - 0.5 * (double)dr) + (double)dr * ((double)Q - col64) / (double)fmaxf(inc, 1e-20f));
        return true;
      }
      col64 = col64 + (double)inc;
      return false;
    }
    if (col + inc > Q) {
      rf = r - 0.5f * dr + dr * (Q - col) / fmaxf(inc, 1e-20f);
      return true;
    }
    col = col + inc;
    return false;
  };
  for (int n = n0; n < n1; n += 2) {
    const float ra = ((float)n + 0.5f) * dr;
    const float rb = ((float)(n + 1) + 0.5f) * dr;
    float na, nb;
    if constexpr (BRICK) {
      // (Skipping the loads of a sample whose bricks repeat the previous sample's, about 40% of consecutive
      // samples, with per-lane predicated loads, measured 8% slower at 512^3; four samples' loads in flight instead
      // of two, 1.35 against 1.20 ms: 102 registers, four waves per SIMD instead of eight.)
      na = sample_brick<OFF>(J.bricks, nx, ny, nz, sx + d[0] * ra, sy + d[1] * ra, sz + d[2] * ra, cs);
      nb = sample_brick<OFF>(J.bricks, nx, ny, nz, sx + d[0] * rb, sy + d[1] * rb, sz + d[2] * rb, cs);
    } else {
      na = sample_r<PAIR, OFF>(J.rho, nx, ny, nz, sx + d[0] * ra, sy + d[1] * ra, sz + d[2] * ra, cs);
      nb = sample_r<PAIR, OFF>(J.rho, nx, ny, nz, sx + d[0] * rb, sy + d[1] * rb, sz + d[2] * rb, cs);
    }
    if (step(ra, na * na * ra * ra * dr)) break;
    if (n + 1 >= n1) break;
    if (step(rb, nb * nb * rb * rb * dr)) break;
  }
  J.rfront[(long long)iu * J.nphi + ip] = rf;
  return rf;
}

// A job's front range: floats [0, 2) hold the least and greatest entry of its front table, floats [2 + 2b, 4 + 2b)
// the least and greatest front radius of the job's workgroup b (256 rays), which the casts write with plain stores
// and rf_range_kernel then reduces into [0, 2). NaN marks a workgroup (or lane) without a ray.
__device__ __forceinline__ void block_rf_range(float rf, float* range, int b) {
  __shared__ float lo_[4], hi_[4];
  float lo = rf, hi = rf;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {   // fminf / fmaxf ignore NaN operands
    lo = fminf(lo, __shfl_xor(lo, off, 64));
    hi = fmaxf(hi, __shfl_xor(hi, off, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    lo_[threadIdx.x >> 6] = lo;
    hi_[threadIdx.x >> 6] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    range[2 + 2 * b] = fminf(fminf(lo_[0], lo_[1]), fminf(lo_[2], lo_[3]));
    range[3 + 2 * b] = fmaxf(fmaxf(hi_[0], hi_[1]), fmaxf(hi_[2], hi_[3]));
  }
}

template <bool PAIR, bool WIDE, class OFF, class SC, bool BRICK>
__global__ __launch_bounds__(256) void cast_rays_kernel(RayJob J, SC cs) {
  cast_one<PAIR, WIDE, OFF, SC, BRICK>(J, cs, wave_patch(blockIdx.x), threadIdx.x & 63);
}

// The ghost-layered brick array of a density field for sample_brick: brick (p, q, r), p in [0, nx+2), q in
// [0, ny+1), r in [0, nz+1), holds (f(i, ja, ka), f(i, ja, kb), f(i, jb, ka), f(i, jb, kb)) with i = clamp(p-1),
// ja = clamp(q-1), jb = clamp(q), ka = clamp(r-1), kb = clamp(r). One thread per brick of a padded z row; the grid
// is (z chunks, padded y, padded x), so neighbouring rows run side by side on all XCDs.
__device__ __forceinline__ void brick_cell(const float* __restrict__ f, float4* __restrict__ B, int nx, int ny, int nz,
                                           int p, int q, int r) {
  const long long i = clampi(p - 1, 0, nx - 1);
  const long long ra = (i * ny + clampi(q - 1, 0, ny - 1)) * nz, rb = (i * ny + clampi(q, 0, ny - 1)) * nz;
  const int ka = clampi(r - 1, 0, nz - 1), kb = clampi(r, 0, nz - 1);
  B[((long long)p * (ny + 1) + q) * (nz + 1) + r] = make_float4(f[ra + ka], f[ra + kb], f[rb + ka], f[rb + kb]);
}

__global__ __launch_bounds__(256) void ray_bricks_kernel(const float* __restrict__ f, float4* __restrict__ B, int nx,
                                                         int ny, int nz) {
  // one padded row (nz + 1 bricks) per block: no block is left with a single brick of a row's tail
  for (int r = threadIdx.x; r <= nz; r += 256) brick_cell(f, B, nx, ny, nz, blockIdx.y, blockIdx.x, r);
}

// The direction of every (u, phi) cell of a ray grid, (d0, d1, d2, 0) at iu nphi + ip, as ray_dir computes it: the
// casts of a batch read it instead of evaluating the direction (a double-precision sine and cosine, a division
// and a square root) once per ray and star.
__global__ __launch_bounds__(256) void ray_dirs_kernel(int nu, int nphi, float4* __restrict__ dirs) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= nu * nphi) return;
  const int iu = t / nphi;
  float d[3];
  ray_dir(iu, t - iu * nphi, nu, nphi, d);
  dirs[t] = make_float4(d[0], d[1], d[2], 0.0f);
}

// blocks of 256 rays each job needs (64-ray direction patches, ray_of_patch)
__host__ __device__ inline long long ray_blocks(int nu, int nphi) {
  return ((long long)((nu + 7) / 8) * ((nphi + 7) / 8) * 64 + 255) / 256;
}

// floats of a job's front range (NvRayDesc.rf_range): the (least, greatest) pair and one pair per workgroup
__host__ __device__ inline long long range_floats(int nu, int nphi) { return 2 + 2 * ray_blocks(nu, nphi); }

// Many ray-cast jobs in one launch: block b belongs to the job s with start[s] <= b < start[s+1]. The cell
// scaling and the FP64 column are per job (block-uniform branches); paired z loads and 32-bit offsets are shared.
template <bool PAIR, class OFF, bool ANY_BRICK>
__global__ __launch_bounds__(256) void cast_rays_batch_kernel(const RayJob* __restrict__ jobs,
                                                              const int* __restrict__ start, int n, int total) {
  const int b = blockIdx.x;
  if (b >= total) return;
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (start[mid] <= b) lo = mid;
    else hi = mid;
  }
  const RayJob& J = jobs[lo];
  const int patch = wave_patch(b - start[lo]), lane = threadIdx.x & 63;
  auto by_scale = [&](auto brick) {
    constexpr bool BR = decltype(brick)::value;
    if (J.inv != 0.0f) {
      ScalePow2 sc;
      sc.inv = J.inv;
      if (J.wide) return cast_one<PAIR, true, OFF, ScalePow2, BR>(J, sc, patch, lane);
      return cast_one<PAIR, false, OFF, ScalePow2, BR>(J, sc, patch, lane);
    }
    ScaleDiv sc;
    sc.rdx = J.rdx;
    if (J.wide) return cast_one<PAIR, true, OFF, ScaleDiv, BR>(J, sc, patch, lane);
    return cast_one<PAIR, false, OFF, ScaleDiv, BR>(J, sc, patch, lane);
  };
  // without brick jobs the kernel holds the direct sampler alone (the brick sampler's registers would lower the
  // occupancy of every job)
  const float rf = ANY_BRICK && J.bricks ? by_scale(std::true_type{}) : by_scale(std::false_type{});
  if (J.rf_range) block_rf_range(rf, J.rf_range, b - start[lo]);   // every workgroup holds rays of one job only
}

// The brick arrays of several grids in one launch: block b belongs to brick job s with bstart[s] <= b <
// bstart[s+1]; each job's blocks take 256 consecutive bricks of its array in storage order (ray_bricks_kernel's
// work, without its one-row-per-block idle threads: the batch's rows are 49-161 bricks long).
struct BrickJob {
  const float* f;
  float4* B;
  int nx, ny, nz, pad;
};

__global__ __launch_bounds__(256) void ray_bricks_batch_kernel(const BrickJob* __restrict__ jobs,
                                                               const int* __restrict__ bstart, int n, int total) {
  const int b = blockIdx.x;
  if (b >= total) return;
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (bstart[mid] <= b) lo = mid;
    else hi = mid;
  }
  const BrickJob& J = jobs[lo];
  const unsigned w = (unsigned)J.nz + 1u, h = (unsigned)J.ny + 1u;   // bricks per padded row, rows per plane
  const unsigned idx = (unsigned)(b - bstart[lo]) * 256u + threadIdx.x;   // < (nx + 2)(ny + 1)(nz + 1) < 2^31
  const unsigned row = idx / w, r = idx - row * w;                         // row = p (ny + 1) + q
  const unsigned p = row / h, q = row - p * h;
  if (p < (unsigned)J.nx + 2u) brick_cell(J.f, J.B, J.nx, J.ny, J.nz, (int)p, (int)q, (int)r);
}

// a / b through the hardware reciprocal (b > 0; within 2 ulp of the quotient)
struct RcpDiv {
  __device__ float operator()(float a, float b) const { return a * __builtin_amdgcn_rcpf(b); }
};

// cell centre seen from a star: r, u = cos(theta), phi in [0, 2pi) (nebula3d.py:284-290). FAST: the hardware
// square root (within 1 ulp of the exact one) and the hardware reciprocal for r2 / r and atan2's quotient.
template <bool FAST = false>
__device__ __forceinline__ void cell_sph(int i, int j, int k, float sx, float sy, float sz, float dx, float* r_o,
                                         float* u_o, float* phi_o) {
  const float r0 = ((float)i + 0.5f) * dx - sx, r1 = ((float)j + 0.5f) * dx - sy, r2 = ((float)k + 0.5f) * dx - sz;
  const float s = (r0 * r0 + r1 * r1) + r2 * r2;
  const float r = fmaxf(FAST ? __builtin_amdgcn_sqrtf(s) : sqrt_rn(s), 1e-6f);
  float phi = FAST ? nvm::atan2f_core(r1, r0, RcpDiv{}) : nvm::atan2f_det(r1, r0);
  if (phi < 0.0f) phi = phi + 2.0f * PI_F;
  *r_o = r;
  *u_o = FAST ? r2 * __builtin_amdgcn_rcpf(r) : r2 / r;
  *phi_o = phi;
}

// phi index wrapped into [0, nphi): ((ip % nphi) + nphi) % nphi, with the common cases as compares (front_at
// and fuv_at only produce ip in [-1, nphi])
__device__ __forceinline__ int wrap_phi(int ip, int nphi) {
  if (ip >= 0 && ip < nphi) return ip;
  if (ip < 0 && ip >= -nphi) return ip + nphi;
  if (ip >= nphi && ip < 2 * nphi) return ip - nphi;
  return ((ip % nphi) + nphi) % nphi;
}

// 1 / (2 pi) as the shared double reciprocal of the float 2*PI_F (correctly rounded, so within the 2^-52 that
// div_r needs): div_r(phi, R2PI) is phi / (2.0f * PI_F) bit for bit
constexpr Recip R2PI = {1.0 / (double)(2.0f * PI_F), 2.0f * PI_F};
constexpr Recip R6 = {1.0 / 6.0, 6.0f};
// the same divisions as float products (the fast ionize pass): 1 / (2 pi) and 1 / 6 rounded to float
constexpr float INV_2PI_F = (float)(1.0 / (double)(2.0f * PI_F));
constexpr float INV_6_F = (float)(1.0 / 6.0);

// row[pa] and row[pb] for pb = pa + 1 or the wrap pb = 0: one 8-byte load of (q, q+1), q = min(pa, nphi-2), when
// nphi >= 2 (row[0] loaded on its own only at the wrap)
__device__ __forceinline__ void phi_pair(const float* __restrict__ row, int pa, int pb, int nphi, float* va,
                                         float* vb) {
  if (nphi < 2) {
    *va = row[pa];
    *vb = row[pb];
    return;
  }
  const int q = min(pa, nphi - 2);
  typedef float f2u __attribute__((ext_vector_type(2), aligned(4)));
  const f2u v = *reinterpret_cast<const f2u*>(row + q);
  *va = pa == q ? v.x : v.y;
  *vb = pb == q + 1 ? v.y : row[pb];
}

// The bilinear front radius at direction (u, phi) (nebula3d.py:291-301, rf_at 273-278) in two halves, so that a
// caller can issue the rfront gathers early and use them later: front_taps issues one 8-byte load per table row,
// (row[q], row[q+1]) with q = min(pa, nphi - 2), and front_lerp combines them. The two phi taps are (pa, pa + 1)
// except at the wrap pa = nphi - 1, where they are (nphi - 1, 0): front_lerp then loads row[0] itself (a rare,
// late load), so front_taps has nothing to wait for.
struct FrontTaps {
  float tu, tp;
  float a0, a1, b0, b1;   // (row[q], row[q+1]) of rows ua and ub
  int ua, ub;             // row offsets (ua nphi, ub nphi)
  bool wrap;
};

template <bool FAST = false>
__device__ __forceinline__ void front_taps(const float* __restrict__ rfront, int nu, int nphi, float u, float phi,
                                           FrontTaps& T) {
  const float fu = (u + 1.0f) * 0.5f * (float)nu - 0.5f;
  const float fp = (FAST ? phi * INV_2PI_F : div_r(phi, R2PI)) * (float)nphi - 0.5f;
  const int iu = (int)floorf(fu), ip = (int)floorf(fp);
  T.tu = fu - (float)iu;
  T.tp = fp - (float)ip;
  T.ua = clampi(iu, 0, nu - 1) * nphi;
  T.ub = clampi(iu + 1, 0, nu - 1) * nphi;
  const int pa = wrap_phi(ip, nphi);
  if (nphi < 2) {   // one phi bin: both taps are row[0]
    T.a0 = T.a1 = rfront[T.ua];
    T.b0 = T.b1 = rfront[T.ub];
    T.wrap = false;
    return;
  }
  const int q = min(pa, nphi - 2);
  typedef float f2u __attribute__((ext_vector_type(2), aligned(4)));
  const f2u va = *reinterpret_cast<const f2u*>(rfront + T.ua + q);
  const f2u vb = *reinterpret_cast<const f2u*>(rfront + T.ub + q);
  T.a0 = va.x;
  T.a1 = va.y;
  T.b0 = vb.x;
  T.b1 = vb.y;
  T.wrap = pa != q;
}

__device__ __forceinline__ float front_lerp(const float* __restrict__ rfront, const FrontTaps& T) {
  float ra0 = T.a0, ra1 = T.a1, rb0 = T.b0, rb1 = T.b1;
  if (T.wrap) {   // taps (nphi - 1, 0): the loaded pair is (nphi - 2, nphi - 1)
    ra0 = T.a1;
    rb0 = T.b1;
    ra1 = rfront[T.ua];
    rb1 = rfront[T.ub];
  }
  const float av = ra0 * (1.0f - T.tp) + ra1 * T.tp;
  const float bv = rb0 * (1.0f - T.tp) + rb1 * T.tp;
  return av * (1.0f - T.tu) + bv * T.tu;
}

// bilinear front radius at direction (u, phi)
__device__ __forceinline__ float front_at(const float* __restrict__ rfront, int nu, int nphi, float u, float phi) {
  FrontTaps T;
  front_taps(rfront, nu, nphi, u, phi, T);
  return front_lerp(rfront, T);
}

// x / d for one divisor and three numerators: the shared reciprocal where d is a positive normal float (the
// identical IEEE quotients, nv_divide.h), IEEE division otherwise (a raw density of 0 or below the float range)
__device__ __forceinline__ void div3(float a, float b, float c, float d, float* qa, float* qb, float* qc) {
  if (d >= 1.17549435e-38f && d <= 3.40282347e+38f) {
    const float m[3] = {a, b, c};
    float v[3];
    div_rn<3>(m, recip(d), v);
    *qa = v[0];
    *qb = v[1];
    *qc = v[2];
  } else {
    *qa = a / d;
    *qb = b / d;
    *qc = c / d;
  }
}

// apply_front (nebula3d.py:281-307)
__global__ __launch_bounds__(256) void apply_front_kernel(const float* __restrict__ rfront, int nu, int nphi,
                                                          float sx, float sy, float sz, float dx, Recip rwidth,
                                                          float* xi, float* dfront, int nx, int ny, int nz) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long long)nx * ny * nz) return;
  const int k = (int)(q % nz);
  const long long ij = q / nz;
  const int j = (int)(ij % ny), i = (int)(ij / ny);
  float r, u, phi;
  cell_sph(i, j, k, sx, sy, sz, dx, &r, &u, &phi);
  const float rf = front_at(rfront, nu, nphi, u, phi);
  const float val = fminf(fmaxf(div_r(rf - r, rwidth) + 0.5f, 0.0f), 1.0f);
  if (val > xi[q]) xi[q] = val;
  dfront[q] = fminf(dfront[q], r - rf);
}

// smooth_xi (nebula3d.py:310-315)
__global__ __launch_bounds__(256) void smooth_xi_kernel(const float* __restrict__ src, float* __restrict__ dst,
                                                        int nx, int ny, int nz) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long long)nx * ny * nz) return;
  const int k = (int)(q % nz);
  const long long ij = q / nz;
  const int j = (int)(ij % ny), i = (int)(ij / ny);
  const float m = div_r(src[cell(clampi(i - 1, 0, nx - 1), j, k, ny, nz)] + src[cell(clampi(i + 1, 0, nx - 1), j, k, ny, nz)] +
                   src[cell(i, clampi(j - 1, 0, ny - 1), k, ny, nz)] + src[cell(i, clampi(j + 1, 0, ny - 1), k, ny, nz)] +
                   src[cell(i, j, clampi(k - 1, 0, nz - 1), ny, nz)] + src[cell(i, j, clampi(k + 1, 0, nz - 1), ny, nz)],
                  R6);
  dst[q] = 0.5f * src[q] + 0.5f * m;
}

__global__ __launch_bounds__(256) void to_half_kernel(const float* __restrict__ src, uint16_t* __restrict__ dst,
                                                      long long n) {
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x)
    dst[q] = f32_to_f16_bits(src[q]);
}

__global__ __launch_bounds__(256) void fill_kernel(float* __restrict__ dst, float v, long long n) {
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x)
    dst[q] = v;
}

// ---------------------------------------------------------------------------------------- fused ionize pass
constexpr int IK = 32, IJ = 16, INT_ = IK * IJ;    // tile: 32 z x 16 y, marching along x
constexpr int XK = IK + 2, XJ = IJ + 2, XN = XK * XJ;
constexpr int XR = 4;                               // ring of unsmoothed-xi planes (i-1, i, i+1 read; i+2 written)
static_assert(INT_ <= XN && XN <= 2 * INT_, "every thread owns one or two halo cells of a plane");

struct IonArgs {
  const float* rfronts;
  const float* star_xyz;
  float* xi;
  uint16_t* df_h;
  float* df_f;
  const void *f0, *f1, *f2, *f3;   // state for the CFL max: (rho, m) float32, or (rho, v) binary16
  float* smax;
  int nu, nphi, nstars, nx, ny, nz, ci;
  int tiles_k, tiles_j, tiles;     // tile grid: z tiles, y tiles, all tiles (x chunks of ci planes slowest)
  float dx, cn2, ci2;
  Recip rwidth;                    // the ramp width and 1 / width, correctly rounded (its shared reciprocal)
  const float* rf_ranges;          // per star (least, greatest) front radius of its table, or null (no culling)
  float rw;                        // 1 / width rounded to float (the fast pass's ramp division)
};

// One 32 z x 16 y tile of the fused ionize pass, marching over x chunk bi.
//
// Per plane i: the unsmoothed xi and dfront of the tile plus a one-cell y/z halo (XN cells: thread t owns cell t,
// and cell t + 512 where that exists) go to an LDS ring; plane i is smoothed from planes i-1, i, i+1. The first
// star's rfront gathers for a thread's cells of plane i+2 are issued before plane i is smoothed and stored, and
// used after: the smoothing, the stores and the CFL max run while the gathers are in flight.
//
// Star culling (with A.rf_ranges and two or more stars). xi is the maximum over stars of the front ramp and dfront
// the minimum of r - rf, so the stars may be visited in any order. A cell visits first the star whose front may
// lie nearest (least r - rf_min), and then skips a star when r - rf >= max(df, w) is certain for it (w the ramp
// width): then its ramp is 0 and min(df, r - rf) is df, so skipping it changes no bit. The test uses a lower
// bound on r (the hardware square root of the same sum cell_sph takes the exact root of, within an ulp) and the
// star's greatest front radius, each widened by 2^-18 relative, far beyond the rounding of r and of the
// interpolated rf. A cell away from every front then evaluates one star's direction and front instead of all.
template <bool SMAX, bool VEL, bool CULL, bool FAST>
__device__ __forceinline__ void ionize_tile(const IonArgs& A, int bk, int bj, int bi) {
  __shared__ float sx_[XR][XN];   // unsmoothed xi (max over stars of the front ramp)
  __shared__ float sd_[XR][XN];   // fp32 dfront (min over stars of r - rf)
  __shared__ float red[INT_ / 64];
  const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * IK + tx;
  const int nx = A.nx, ny = A.ny, nz = A.nz;
  const int k0 = bk * IK, j0 = bj * IJ, i0 = bi * A.ci;
  const int i1 = min(i0 + A.ci, nx);
  const int k = k0 + tx, j = j0 + ty;
  const bool active = k < nz && j < ny;
  const long long tab = (long long)A.nu * A.nphi;
  // the star table is read-only for the whole launch: through the constant address space its loads (at a
  // wave-uniform index) are scalar, instead of a vector load and wait per star and cell
  const CONST_AS float* stars = (const CONST_AS float*)A.star_xyz;

  // CULL: some grid of the launch carries front ranges; this tile culls with two to 64 stars
  const bool cull = CULL && A.rf_ranges != nullptr && A.nstars >= 2 && A.nstars <= 64;
  __shared__ float ranges[CULL ? 128 : 1];   // per star its least and greatest front radius
  if (cull && tid < 2 * A.nstars) ranges[tid] = A.rf_ranges[(tid >> 1) * range_floats(A.nu, A.nphi) + (tid & 1)];
  if (cull) __syncthreads();
  // halo-plane cell t of plane p (clamped cell positions, as ldc in smooth_xi): its distance, direction and the
  // front taps of star st
  auto cell_of = [&](int p, int t, int& ic, int& jc, int& kc) {
    const int jj = t / XK, kk = t - jj * XK;
    ic = clampi(p, 0, nx - 1);
    jc = clampi(j0 + jj - 1, 0, ny - 1);
    kc = clampi(k0 + kk - 1, 0, nz - 1);
  };
  auto star_taps = [&](int p, int t, int st, float& r, FrontTaps& T) {
    int ic, jc, kc;
    cell_of(p, t, ic, jc, kc);
    float u, phi;
    cell_sph<FAST>(ic, jc, kc, stars[3 * st], stars[3 * st + 1], stars[3 * st + 2], A.dx, &r, &u, &phi);
    front_taps<FAST>(A.rfronts + st * tab, A.nu, A.nphi, u, phi, T);
  };
  // the distance from star st to the cell centre by the hardware square root of the sum cell_sph forms (within an
  // ulp of its correctly rounded root; with FAST, the bits cell_sph takes before its floor of 1e-6)
  auto r_fast = [&](int ic, int jc, int kc, int st) {
    const float r0 = ((float)ic + 0.5f) * A.dx - stars[3 * st];
    const float r1 = ((float)jc + 0.5f) * A.dx - stars[3 * st + 1];
    const float r2 = ((float)kc + 0.5f) * A.dx - stars[3 * st + 2];
    return __builtin_amdgcn_sqrtf((r0 * r0 + r1 * r1) + r2 * r2);
  };
  // the star this cell visits first: 0, or with culling the one of least r - rf_min
  auto first_star = [&](int p, int t) {
    if (!cull) return 0;
    int ic, jc, kc;
    cell_of(p, t, ic, jc, kc);
    int best = 0;
    float bv = r_fast(ic, jc, kc, 0) - ranges[0];
    for (int st = 1; st < A.nstars; ++st) {
      const float v = r_fast(ic, jc, kc, st) - ranges[2 * st];
      if (v < bv) {
        bv = v;
        best = st;
      }
    }
    return best;
  };
  // true when star st certainly leaves (x, df) of the cell unchanged: r - rf >= max(df, w)
  auto skip_star = [&](int ic, int jc, int kc, int st, float df) {
    const float m = fmaxf(df, A.rwidth.b) + ranges[2 * st + 1];
    return r_fast(ic, jc, kc, st) * (1.0f - 0x1p-18f) >= m * (1.0f + 0x1p-18f);
  };
  // one star's ramp into the running max (xi) and min (dfront)
  auto accumulate = [&](float r, float rf, float& x, float& df) {
    const float val = fminf(fmaxf((FAST ? (rf - r) * A.rw : div_r(rf - r, A.rwidth)) + 0.5f, 0.0f), 1.0f);
    if (val > x) x = val;
    df = fminf(df, r - rf);
  };
  // every star but s0 of halo cell t of plane p, written to the ring; x, df hold star s0's contribution
  auto finish_cell = [&](int p, int t, int s0, float x, float df) {
    int ic, jc, kc;
    cell_of(p, t, ic, jc, kc);
    for (int st = 0; st < A.nstars; ++st) {
      if (st == s0 || (cull && skip_star(ic, jc, kc, st, df))) continue;
      float r;
      FrontTaps T;
      star_taps(p, t, st, r, T);
      accumulate(r, front_lerp(A.rfronts + st * tab, T), x, df);
    }
    const int sl = ((p % XR) + XR) % XR;
    sx_[sl][t] = x;
    sd_[sl][t] = df;
  };
  // whole plane p: every halo cell, every star (xi.zero_(), df.fill_(1e3): sim.py:930-931)
  auto make_plane = [&](int p) {
    for (int t = tid; t < XN; t += INT_) {
      float x = 0.0f, df = 1e3f;
      int s0 = -1;
      if (A.nstars > 0) {
        s0 = first_star(p, t);
        float r;
        FrontTaps T;
        star_taps(p, t, s0, r, T);
        accumulate(r, front_lerp(A.rfronts + s0 * tab, T), x, df);
      }
      finish_cell(p, t, s0, x, df);
    }
  };

  make_plane(i0 - 1);
  make_plane(i0);
  make_plane(i0 + 1);
  __syncthreads();
  float tmax = 0.0f;
  for (int i = i0; i < i1; ++i) {
    // the state of plane i for the CFL max, then the first star's front gathers of this thread's cells of plane i+2
    const long long o = cell(i, j, k, ny, nz);
    float f0 = 0.0f, f1 = 0.0f, f2 = 0.0f, f3 = 0.0f;
    if (SMAX && active) {
      if constexpr (VEL) {
        f1 = ldf((const uint16_t*)A.f1, o);
        f2 = ldf((const uint16_t*)A.f2, o);
        f3 = ldf((const uint16_t*)A.f3, o);
      } else {
        f0 = ((const float*)A.f0)[o];
        f1 = ((const float*)A.f1)[o];
        f2 = ((const float*)A.f2)[o];
        f3 = ((const float*)A.f3)[o];
      }
    }
    const bool next = i + 2 <= i1;   // plane i1 is the last one a smoothing of this chunk reads
    const bool second = tid + INT_ < XN;   // this thread also owns a cell of the second pass
    float r0 = 0.0f, r1 = 0.0f;
    FrontTaps T0, T1;
    int s0 = -1, s1 = -1;
    if (next && A.nstars > 0) {
      s0 = first_star(i + 2, tid);
      star_taps(i + 2, tid, s0, r0, T0);
      if (second) {
        s1 = first_star(i + 2, tid + INT_);
        star_taps(i + 2, tid + INT_, s1, r1, T1);
      }
    }

    if (active) {
      const int c = (ty + 1) * XK + tx + 1;
      const float* pc = sx_[((i % XR) + XR) % XR];
      const float* pm = sx_[(((i - 1) % XR) + XR) % XR];
      const float* pp = sx_[(((i + 1) % XR) + XR) % XR];
      // neighbours in the order of nebula3d.py:313-314
      const float sum = pm[c] + pp[c] + pc[c - XK] + pc[c + XK] + pc[c - 1] + pc[c + 1];
      const float m = FAST ? sum * INV_6_F : div_r(sum, R6);
      const float xs = 0.5f * pc[c] + 0.5f * m;
      A.xi[o] = xs;
      const float df = sd_[((i % XR) + XR) % XR][c];
      if (A.df_h) A.df_h[o] = f32_to_f16_bits(df);
      else A.df_f[o] = df;
      if constexpr (SMAX) {
        float s;
        const float c2 = A.cn2 + xs * (A.ci2 - A.cn2);
        const float c = FAST ? __builtin_amdgcn_sqrtf(c2) : sqrt_rn(c2);
        if constexpr (VEL) {
          s = ((fabsf(f1) + fabsf(f2)) + fabsf(f3)) + c;
        } else {
          float v0, v1, v2;
          if constexpr (FAST) {
            const float ir = __builtin_amdgcn_rcpf(f0);
            v0 = f1 * ir;
            v1 = f2 * ir;
            v2 = f3 * ir;
          } else {
            div3(f1, f2, f3, f0, &v0, &v1, &v2);
          }
          s = ((fabsf(v0) + fabsf(v1)) + fabsf(v2)) + c;
        }
        tmax = fmaxf(tmax, s);
      }
    }

    if (next) {
      // plane i+2 goes to the ring slot of plane i-2, which no thread reads after the last barrier
      float x = 0.0f, df = 1e3f;
      if (A.nstars > 0) accumulate(r0, front_lerp(A.rfronts + s0 * tab, T0), x, df);
      finish_cell(i + 2, tid, s0, x, df);
      if (second) {
        x = 0.0f;
        df = 1e3f;
        if (A.nstars > 0) accumulate(r1, front_lerp(A.rfronts + s1 * tab, T1), x, df);
        finish_cell(i + 2, tid + INT_, s1, x, df);
      }
    }
    __syncthreads();
  }
  if constexpr (SMAX) block_max_atomic(tmax, red, A.smax);
}

// One tile per workgroup, each XCD a contiguous run of tiles (z fastest, then y, then the x chunk): its cells then
// see a narrower cone of directions from each star, so the front-table rows they gather stay in the XCD's L2 (the
// 512^3 wall's pass: 1.07 against 1.15 ms with the tiles dealt round robin; the same bits).
template <bool SMAX, bool VEL, bool FAST>
__global__ __launch_bounds__(INT_) void ionize_apply_kernel(IonArgs A, int per_xcd) {
  const int t = xcd_tile(per_xcd);
  if (t >= A.tiles) return;
  ionize_tile<SMAX, VEL, false, FAST>(A, t % A.tiles_k, (t / A.tiles_k) % A.tiles_j, t / (A.tiles_k * A.tiles_j));
}

// Many grids in one launch: tile t of the flat list belongs to member s with start[s] <= t < start[s+1]. XCD x takes
// the contiguous run of tiles [xoff[x], xoff[x+1]) (workgroup id b runs on XCD b % 8 and takes tile xoff[b % 8] +
// b / 8), cut by the plan so that the runs carry about equal work: each member's tiles then run on one XCD, or on
// two adjacent ones, and its front tables (2 MB per star at 512 x 1024 rays) are gathered into that XCD's L2 only.
// On the 64-member batch (160 stars, star culling): 1.51-1.52 ms against 1.55-1.56 with the tiles dealt round robin,
// where every XCD gathers from every member's tables and half the pass's L2 lookups missed. Equal tile counts per
// XCD, without the work weights, measured 2.34 against 2.04 ms before the culling: members differ in star count.
template <bool SMAX, bool VEL, bool CULL, bool FAST>
__global__ __launch_bounds__(INT_) void ionize_apply_batch_kernel(const IonArgs* __restrict__ args,
                                                                  const int* __restrict__ start, int n,
                                                                  const int* __restrict__ xoff) {
  const int x = blockIdx.x & 7;
  const int t = xoff[x] + (int)(blockIdx.x >> 3);
  if (t >= xoff[x + 1]) return;
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (start[mid] <= t) lo = mid;
    else hi = mid;
  }
  const IonArgs& A = args[lo];
  const int l = t - start[lo];
  const int bk = l % A.tiles_k, bj = (l / A.tiles_k) % A.tiles_j, bi = l / (A.tiles_k * A.tiles_j);
  ionize_tile<SMAX, VEL, CULL, FAST>(A, bk, bj, bi);
}

// ---------------------------------------------------------------------------------------- FUV
// cast_fuv (nebula3d.py:320-352)
template <bool PAIR, class OFF, class SC>
__global__ __launch_bounds__(256) void cast_fuv_kernel(const float* __restrict__ rho, const float* __restrict__ xi,
                                                       int nx, int ny, int nz, float sx, float sy, float sz,
                                                       float kappa, float dr, int nsteps, SC cs, float binr,
                                                       float bx, float by, float bz, float* __restrict__ att, int nr,
                                                       int nu, int nphi) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nang = (long long)nu * nphi;
  if (t >= nang) return;
  const int iu = (int)(t / nphi);
  float d[3];
  ray_dir(iu, (int)(t - (long long)iu * nphi), nu, nphi, d);
  const float st[3] = {sx, sy, sz}, bm[3] = {bx, by, bz};
  float t0, t1;
  box_range(st, d, bm, &t0, &t1);
  const int n0 = (int)(t0 / dr);
  const int n1 = min(nsteps, (int)(t1 / dr) + 2);
  float tau = 0.0f;
  int ib = 0;
  for (int n = n0; n < n1; ++n) {
    const float r = ((float)n + 0.5f) * dr;
    const float qx = sx + d[0] * r, qy = sy + d[1] * r, qz = sz + d[2] * r;
    const float n_tot = sample_r<PAIR, OFF>(rho, nx, ny, nz, qx, qy, qz, cs);
    const float x_i = sample_r<PAIR, OFF>(xi, nx, ny, nz, qx, qy, qz, cs);
    tau = tau + kappa * n_tot * (1.0f - 0.92f * x_i) * dr;
    const int nb = (int)(r / binr);
    if (nb > ib) {
      const float e = nvm::expf_det(-tau);
      for (int b = ib; b < min(nb, nr); ++b) att[(long long)b * nang + t] = e;
      ib = nb;
    }
    if (ib >= nr) break;
  }
  const float e = nvm::expf_det(-tau);
  for (int b = ib; b < nr; ++b) att[(long long)b * nang + t] = e;
}

// trilinear att at (r/binr - 1/2, u, phi) (nebula3d.py:362-387); rbinr = 1/binr as the shared reciprocal
__device__ __forceinline__ float fuv_at(const float* __restrict__ att, int nr, int nu, int nphi, float r, float u,
                                        float phi, const Recip& rbinr) {
  const float fr = div_r(r, rbinr) - 0.5f;
  const float fu = (u + 1.0f) * 0.5f * (float)nu - 0.5f;
  const float fp = div_r(phi, R2PI) * (float)nphi - 0.5f;
  const int ir = (int)floorf(fr), iu = (int)floorf(fu), ip = (int)floorf(fp);
  const float tr = fr - (float)ir, tu = fu - (float)iu, tp = fp - (float)ip;
  const int pa = wrap_phi(ip, nphi), pb = wrap_phi(ip + 1, nphi);
  float acc = 0.0f;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int rr = clampi(ir + a, 0, nr - 1), uu = clampi(iu + b, 0, nu - 1);
      float v0, v1;
      phi_pair(att + ((long long)rr * nu + uu) * nphi, pa, pb, nphi, &v0, &v1);
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const float w = (a == 0 ? 1.0f - tr : tr) * (b == 0 ? 1.0f - tu : tu) * (c == 0 ? 1.0f - tp : tp);
        acc = acc + w * (c == 0 ? v0 : v1);
      }
    }
  return acc;
}

// apply_fuv (nebula3d.py:355-388)
__global__ __launch_bounds__(256) void apply_fuv_kernel(const float* __restrict__ att, int nr, int nu, int nphi,
                                                        float sx, float sy, float sz, float lum, float dx,
                                                        Recip rbinr, float* g0, int nx, int ny, int nz) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long long)nx * ny * nz) return;
  const int k = (int)(q % nz);
  const long long ij = q / nz;
  const int j = (int)(ij % ny), i = (int)(ij / ny);
  float r, u, phi;
  cell_sph(i, j, k, sx, sy, sz, dx, &r, &u, &phi);
  const float acc = fuv_at(att, nr, nu, nphi, r, u, phi, rbinr);
  const float r2 = fmaxf(r * r, 0.0004f);
  g0[q] = g0[q] + lum * acc / r2;
}

// the per-cell part of Sim.fuv_field (sim.py:955-978), all stars in order, one store
__global__ __launch_bounds__(256) void fuv_apply_kernel(const float* __restrict__ atts, int nr, int nu, int nphi,
                                                        const float* __restrict__ star_xyz,
                                                        const float* __restrict__ lums, int nstars, float dx,
                                                        Recip rbinr, uint16_t* g0_h, float* g0_f, int nx, int ny,
                                                        int nz) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long long)nx * ny * nz) return;
  const int k = (int)(q % nz);
  const long long ij = q / nz;
  const int j = (int)(ij % ny), i = (int)(ij / ny);
  const long long tab = (long long)nr * nu * nphi;
  float g = 0.0f;   // acc.zero_() (sim.py:957)
  for (int s = 0; s < nstars; ++s) {
    float r, u, phi;
    cell_sph(i, j, k, star_xyz[3 * s], star_xyz[3 * s + 1], star_xyz[3 * s + 2], dx, &r, &u, &phi);
    const float acc = fuv_at(atts + s * tab, nr, nu, nphi, r, u, phi, rbinr);
    const float r2 = fmaxf(r * r, 0.0004f);
    g = g + lums[s] * acc / r2;
  }
  if (g0_h) g0_h[q] = f32_to_f16_bits(g);
  else g0_f[q] = g;
}

inline int blocks_for(long long n, int bs) { return (int)((n + bs - 1) / bs); }

}  // namespace nv

using namespace nv;

// Calls f(pair, offset, scale) with the sampler's compile-time forms for this grid: paired z loads when nz >= 2,
// 32-bit byte offsets (and 24-bit index products) up to 2^30 cells with nx ny < 2^24, and the power-of-two or
// shared-reciprocal cell scaling. cells: the grid's cells times the element size in floats.
template <class F>
static void dispatch_sampler(int nz, long long cells, long long rows, float dx, F&& f) {
  const float inv = pow2_reciprocal(dx);
  ScaleDiv sd;
  sd.rdx = Recip{1.0 / (double)dx, dx};   // correctly rounded: |r dx - 1| <= 2^-53, inside the bound div_r needs
  ScalePow2 sp;
  sp.inv = inv;
  auto by_scale = [&](auto pair, auto off) {
    if (inv != 0.0f) f(pair, off, sp);
    else f(pair, off, sd);
  };
  auto by_off = [&](auto pair) {
    if (cells <= (1LL << 30) && rows < (1LL << 24)) by_scale(pair, uint32_t{});
    else by_scale(pair, uint64_t{});
  };
  if (nz >= 2) by_off(std::true_type{});
  else by_off(std::false_type{});
}

static void fill_ray_job(const float* rho, int nx, int ny, int nz, float sx, float sy, float sz, float Q, float dr,
                         int nsteps, float dx, float rcap, float bx, float by, float bz, float* rfront, int nu,
                         int nphi, bool wide, RayJob& J) {
  J.rho = rho; J.bricks = nullptr; J.dirs = nullptr; J.rf_range = nullptr; J.rfront = rfront;
  J.rdx = Recip{1.0 / (double)dx, dx};
  J.nx = nx; J.ny = ny; J.nz = nz; J.nsteps = nsteps; J.nu = nu; J.nphi = nphi; J.wide = wide ? 1 : 0;
  J.sx = sx; J.sy = sy; J.sz = sz; J.Q = Q; J.dr = dr; J.dx = dx; J.rcap = rcap; J.bx = bx; J.by = by; J.bz = bz;
  J.inv = pow2_reciprocal(dx);
  J.inside = (sx > 0.0f && sx < bx && sy > 0.0f && sy < by && sz > 0.0f && sz < bz) ? 1 : 0;
}

extern "C" long long nv_ray_bricks_size(int nx, int ny, int nz);

static int cast_rays_launch(const char* name, bool wide, const float* rho, const float* bricks, int nx, int ny,
                            int nz, float sx, float sy, float sz, float Q, float dr, int nsteps, float dx, float rcap,
                            float bx, float by, float bz, float* rfront, int nu, int nphi, void* stream) {
  clear_error();
  NV_REQUIRE(rho && rfront, "%s: null pointer", name);
  NV_REQUIRE(nx > 0 && ny > 0 && nz > 0 && nu > 0 && nphi > 0, "%s: bad sizes", name);
  NV_REQUIRE(dr > 0.0f && dx > 0.0f, "%s: dr and dx must be > 0", name);
  NV_REQUIRE((long long)nx * ny * nz < (1LL << 32), "%s: grids above 2^32 cells are not supported", name);
  NV_REQUIRE(!bricks || ((uintptr_t)bricks & 15) == 0, "%s: the brick array must be 16-byte aligned", name);
  NV_REQUIRE(!bricks || nv_ray_bricks_size(nx, ny, nz) / 16 < (1LL << 32), "%s: brick grid above 2^32 bricks", name);
  RayJob J;
  fill_ray_job(rho, nx, ny, nz, sx, sy, sz, Q, dr, nsteps, dx, rcap, bx, by, bz, rfront, nu, nphi, wide, J);
  J.bricks = reinterpret_cast<const float4*>(bricks);
  const long long blocks = ray_blocks(nu, nphi);
  const long long cells = (long long)nx * ny * nz;
  if (bricks) {
    // 16-byte bricks of the padded grid: 32-bit byte offsets up to 2^28 bricks, 24-bit products below 2^24 rows
    const long long prow = (long long)(nx + 2) * (ny + 1);
    dispatch_sampler(2, prow * (nz + 1) * 4, prow, dx, [&](auto, auto off, auto sc) {
      using O = decltype(off);
      using SC = decltype(sc);
      auto kern = wide ? cast_rays_kernel<true, true, O, SC, true> : cast_rays_kernel<true, false, O, SC, true>;
      hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(256), 0, as_stream(stream), J, sc);
    });
  } else {
    dispatch_sampler(nz, cells, (long long)nx * ny, dx, [&](auto pair, auto off, auto sc) {
      constexpr bool P = decltype(pair)::value;
      using O = decltype(off);
      using SC = decltype(sc);
      auto kern = wide ? cast_rays_kernel<P, true, O, SC, false> : cast_rays_kernel<P, false, O, SC, false>;
      hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(256), 0, as_stream(stream), J, sc);
    });
  }
  return check_launch(name);
}

extern "C" int nv_cast_rays(const float* rho, int nx, int ny, int nz, float sx, float sy, float sz, float Q, float dr,
                            int nsteps, float dx, float rcap, float bx, float by, float bz, float* rfront, int nu,
                            int nphi, void* stream) {
  return cast_rays_launch("nv_cast_rays", false, rho, nullptr, nx, ny, nz, sx, sy, sz, Q, dr, nsteps, dx, rcap, bx,
                          by, bz, rfront, nu, nphi, stream);
}

extern "C" int nv_cast_rays_wide(const float* rho, int nx, int ny, int nz, float sx, float sy, float sz, float Q,
                                 float dr, int nsteps, float dx, float rcap, float bx, float by, float bz,
                                 float* rfront, int nu, int nphi, void* stream) {
  return cast_rays_launch("nv_cast_rays_wide", true, rho, nullptr, nx, ny, nz, sx, sy, sz, Q, dr, nsteps, dx, rcap,
                          bx, by, bz, rfront, nu, nphi, stream);
}

extern "C" long long nv_ray_bricks_size(int nx, int ny, int nz) {
  if (nx <= 0 || ny <= 0 || nz <= 0) return -1;
  return 16LL * (nx + 2) * (ny + 1) * (nz + 1);
}

extern "C" int nv_ray_bricks(const float* rho, int nx, int ny, int nz, float* bricks, void* stream) {
  clear_error();
  NV_REQUIRE(rho && bricks, "nv_ray_bricks: null pointer");
  NV_REQUIRE(((uintptr_t)bricks & 15) == 0, "nv_ray_bricks: the brick array must be 16-byte aligned");
  NV_REQUIRE(nx > 0 && ny > 0 && nz > 0 && nx <= 65533 && ny <= 65534, "nv_ray_bricks: bad sizes");
  const long long cells = (long long)nx * ny * nz;
  const char* r = reinterpret_cast<const char*>(rho);
  const char* b = reinterpret_cast<const char*>(bricks);
  NV_REQUIRE(b + nv_ray_bricks_size(nx, ny, nz) <= r || r + 4 * cells <= b,
             "nv_ray_bricks: the brick array overlaps the field");
  hipLaunchKernelGGL(ray_bricks_kernel, dim3((unsigned)(ny + 1), (unsigned)(nx + 2)), dim3(256), 0, as_stream(stream),
                     rho, reinterpret_cast<float4*>(bricks), nx, ny, nz);
  return check_launch("nv_ray_bricks");
}

extern "C" long long nv_ray_range_floats(int nu, int nphi) {
  return nu > 0 && nphi > 0 ? range_floats(nu, nphi) : -1;
}

extern "C" int nv_ray_dirs(int nu, int nphi, float* dirs, void* stream) {
  clear_error();
  NV_REQUIRE(dirs && ((uintptr_t)dirs & 15) == 0, "nv_ray_dirs: dirs must be a 16-byte aligned buffer");
  NV_REQUIRE(nu > 0 && nphi > 0 && (long long)nu * nphi < (1LL << 31), "nv_ray_dirs: bad sizes");
  hipLaunchKernelGGL(ray_dirs_kernel, dim3((unsigned)(((long long)nu * nphi + 255) / 256)), dim3(256), 0,
                     as_stream(stream), nu, nphi, reinterpret_cast<float4*>(dirs));
  return check_launch("nv_ray_dirs");
}

extern "C" int nv_cast_rays_ex(const float* rho, const float* bricks, int nx, int ny, int nz, float sx, float sy,
                               float sz, float Q, float dr, int nsteps, float dx, float rcap, float bx, float by,
                               float bz, float* rfront, int nu, int nphi, int wide, void* stream) {
  return cast_rays_launch("nv_cast_rays_ex", wide != 0, rho, bricks, nx, ny, nz, sx, sy, sz, Q, dr, nsteps, dx, rcap,
                          bx, by, bz, rfront, nu, nphi, stream);
}

// ------------------------------------------------------------------------------------------ no-allocation cast
// nv_cast_rays_noalloc builds the density's brick array in `scratch`, a static device array of this library (one
// per GPU, held until the process ends), instead of a caller buffer: for callers that cannot allocate device memory.
// 128 MiB holds the bricks of grids up to about 200^3 (a 160^3 grid needs 64.1 MiB). Host code never passes the
// array's name to a kernel: each device's address comes from hipGetSymbolAddress on the first call with that device
// current, and is kept per device. Every brick the casts read is written by the same call's brick build, so nothing
// needs clearing; two calls that overlap on one GPU (other streams or threads) share the array and race.
namespace nv {
constexpr long long kNoAllocBrickBytes = 128LL << 20;
static __device__ __align__(256) float4 scratch[kNoAllocBrickBytes / 16];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static float4* g_scratch_ptr[kMaxDevices];

// this device's address of `scratch`, looked up once per device (with the device current)
static int scratch_of_current_device(const char* name, float4** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    set_error("%s: no current device (or device id %d above %d)", name, dev, kMaxDevices - 1);
    return NV_ERR_LAUNCH;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error("%s: hipGetSymbolAddress(scratch) on device %d failed: %s", name, dev, hipGetErrorString(e));
      return NV_ERR_LAUNCH;
    }
    g_scratch_ptr[dev] = static_cast<float4*>(p);
  }
  *out = g_scratch_ptr[dev];
  return NV_OK;
}
}  // namespace nv

extern "C" int nv_cast_rays_noalloc(const float* rho, int nx, int ny, int nz, float sx, float sy, float sz, float Q,
                                    float dr, int nsteps, float dx, float rcap, float bx, float by, float bz,
                                    float* rfront, int nu, int nphi, int wide, void* stream) {
  using namespace nv;
  clear_error();
  NV_REQUIRE(rho && rfront, "nv_cast_rays_noalloc: null pointer");
  NV_REQUIRE(nx > 0 && ny > 0 && nz > 0 && nu > 0 && nphi > 0, "nv_cast_rays_noalloc: bad sizes");
  const long long need = nv_ray_bricks_size(nx, ny, nz);
  NV_REQUIRE(need <= kNoAllocBrickBytes,
             "nv_cast_rays_noalloc: the %dx%dx%d grid's brick array needs %lld bytes, above the library's %lld-byte "
             "static scratch (nv_ray_bricks + nv_cast_rays_ex take a caller buffer of any size)",
             nx, ny, nz, need, kNoAllocBrickBytes);
  float4* B = nullptr;
  const int rc = scratch_of_current_device("nv_cast_rays_noalloc", &B);
  if (rc != NV_OK) return rc;
  hipLaunchKernelGGL(ray_bricks_kernel, dim3((unsigned)(ny + 1), (unsigned)(nx + 2)), dim3(256), 0, as_stream(stream),
                     rho, B, nx, ny, nz);
  if (check_launch("nv_cast_rays_noalloc") != NV_OK) return NV_ERR_LAUNCH;
  return cast_rays_launch("nv_cast_rays_noalloc", wide != 0, rho, reinterpret_cast<const float*>(B), nx, ny, nz, sx,
                          sy, sz, Q, dr, nsteps, dx, rcap, bx, by, bz, rfront, nu, nphi, stream);
}

extern "C" int nv_apply_front(const float* rfront, int nu, int nphi, float sx, float sy, float sz, float dx,
                              float width, float* xi, float* dfront, int nx, int ny, int nz, void* stream) {
  clear_error();
  NV_REQUIRE(rfront && xi && dfront, "nv_apply_front: null pointer");
  NV_REQUIRE(nx > 0 && ny > 0 && nz > 0 && nu > 0 && nphi > 0, "nv_apply_front: bad sizes");
  NV_REQUIRE(width >= 1.17549435e-38f && width <= 3.40282347e+38f, "nv_apply_front: width must be a positive normal float");
  const long long n = (long long)nx * ny * nz;
  hipLaunchKernelGGL(apply_front_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, as_stream(stream), rfront, nu, nphi,
                     sx, sy, sz, dx, Recip{1.0 / (double)width, width}, xi, dfront, nx, ny, nz);
  return check_launch("nv_apply_front");
}

extern "C" int nv_smooth_xi(const float* src, float* dst, int nx, int ny, int nz, void* stream) {
  clear_error();
  NV_REQUIRE(src && dst && src != dst, "nv_smooth_xi: null or aliased pointers");
  NV_REQUIRE(nx > 0 && ny > 0 && nz > 0, "nv_smooth_xi: bad sizes");
  const long long n = (long long)nx * ny * nz;
  hipLaunchKernelGGL(smooth_xi_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, as_stream(stream), src, dst, nx, ny,
                     nz);
  return check_launch("nv_smooth_xi");
}

extern "C" int nv_to_half(const float* src, uint16_t* dst, long long n, void* stream) {
  clear_error();
  NV_REQUIRE(src && dst && n >= 0, "nv_to_half: bad arguments");
  if (n == 0) return NV_OK;
  const int blocks = (int)std::min<long long>(blocks_for(n, 256), 16384);
  hipLaunchKernelGGL(to_half_kernel, dim3(blocks), dim3(256), 0, as_stream(stream), src, dst, n);
  return check_launch("nv_to_half");
}

extern "C" int nv_fill_f32(float* dst, float v, long long n, void* stream) {
  clear_error();
  NV_REQUIRE(dst && n >= 0, "nv_fill_f32: bad arguments");
  if (n == 0) return NV_OK;
  const int blocks = (int)std::min<long long>(blocks_for(n, 256), 16384);
  hipLaunchKernelGGL(fill_kernel, dim3(blocks), dim3(256), 0, as_stream(stream), dst, v, n);
  return check_launch("nv_fill_f32");
}

// Planes an ionize tile marches at most: a tile recomputes the unsmoothed xi of two planes beyond its chunk, and
// enough tiles keep every CU busy to the end. The 512^3 wall's pass is fastest with chunks of 64 (1.00 ms against
// 1.01 at 32 and 1.05 at 16). The 64-member batch is fastest with 16 (1.37-1.39 ms against 1.48-1.49 at 64, 1.40 at
// 32 and 12, 1.42 at 8): more, shorter tiles even out the runs of its XCDs. Neither choice changes a bit.
constexpr int kIonDepth = 64;
constexpr int kIonBatchDepth = 16;

static int fill_ionize(const char* name, const float* rfronts, int nu, int nphi, const float* star_xyz, int nstars,
                       float dx, float width, float* xi, uint16_t* dfront_h, float* dfront_f, int nx, int ny, int nz,
                       const void* f0, const void* f1, const void* f2, const void* f3, float cn2, float ci2,
                       float* smax_out, long long jk_total, int max_ci, IonArgs& A) {
  NV_REQUIRE(xi && (nstars == 0 || (rfronts && star_xyz)), "%s: null pointer", name);
  NV_REQUIRE((dfront_h != nullptr) != (dfront_f != nullptr), "%s: pass exactly one dfront buffer", name);
  NV_REQUIRE(nx > 0 && ny > 0 && nz > 0 && nu > 0 && nphi > 0 && nstars >= 0, "%s: bad sizes", name);
  NV_REQUIRE(!smax_out || (f0 && f1 && f2 && f3), "%s: smax needs the density and momentum fields", name);
  NV_REQUIRE(width >= 1.17549435e-38f && width <= 3.40282347e+38f, "%s: width must be a positive normal float", name);
  A.rfronts = rfronts; A.star_xyz = star_xyz; A.xi = xi; A.df_h = dfront_h; A.df_f = dfront_f;
  A.f0 = f0; A.f1 = f1; A.f2 = f2; A.f3 = f3; A.smax = smax_out;
  A.nu = nu; A.nphi = nphi; A.nstars = nstars; A.nx = nx; A.ny = ny; A.nz = nz;
  A.dx = dx; A.cn2 = cn2; A.ci2 = ci2;
  A.rwidth = Recip{1.0 / (double)width, width};
  A.rw = 1.0f / width;
  A.rf_ranges = nullptr;
  A.ci = (int)std::min<long long>(max_ci, std::max<long long>(4, (long long)nx * jk_total / 2048));
  A.tiles_k = (nz + IK - 1) / IK;
  A.tiles_j = (ny + IJ - 1) / IJ;
  A.tiles = A.tiles_k * A.tiles_j * ((nx + A.ci - 1) / A.ci);
  return NV_OK;
}

static long long ion_jk(int ny, int nz) { return (long long)((nz + IK - 1) / IK) * ((ny + IJ - 1) / IJ); }

// the kernel instance of a form (NV_ION_VEL, NV_ION_FAST bits) with or without the CFL maximum
template <bool FAST>
static auto ionize_kernel_of(bool smax, bool vel) {
  return !smax ? ionize_apply_kernel<false, false, FAST>
               : (vel ? ionize_apply_kernel<true, true, FAST> : ionize_apply_kernel<true, false, FAST>);
}

static int ionize_apply_launch(const char* name, int form, const float* rfronts, int nu, int nphi,
                               const float* star_xyz, int nstars, float dx, float width, float* xi, uint16_t* dfront_h,
                               float* dfront_f, int nx, int ny, int nz, const void* f0, const void* f1,
                               const void* f2, const void* f3, float cn2, float ci2, float* smax_out, void* stream) {
  clear_error();
  NV_REQUIRE(form >= 0 && form <= (NV_ION_VEL | NV_ION_FAST), "%s: bad form %d", name, form);
  IonArgs A;
  const int rc = fill_ionize(name, rfronts, nu, nphi, star_xyz, nstars, dx, width, xi, dfront_h, dfront_f, nx, ny, nz,
                             f0, f1, f2, f3, cn2, ci2, smax_out, ion_jk(ny, nz), kIonDepth, A);
  if (rc != NV_OK) return rc;
  const int per_xcd = (A.tiles + 7) / 8;
  const bool vel = (form & NV_ION_VEL) != 0;
  auto kern = (form & NV_ION_FAST) ? ionize_kernel_of<true>(smax_out != nullptr, vel)
                                   : ionize_kernel_of<false>(smax_out != nullptr, vel);
  hipLaunchKernelGGL(kern, dim3(8 * per_xcd), dim3(IK, IJ), 0, as_stream(stream), A, per_xcd);
  return check_launch(name);
}

extern "C" int nv_ionize_apply(const float* rfronts, int nu, int nphi, const float* star_xyz, int nstars, float dx,
                               float width, float* xi, uint16_t* dfront_h, float* dfront_f, int nx, int ny, int nz,
                               const float* rho, const float* mx, const float* my, const float* mz, float cn2,
                               float ci2, float* smax_out, void* stream) {
  return ionize_apply_launch("nv_ionize_apply", 0, rfronts, nu, nphi, star_xyz, nstars, dx, width, xi, dfront_h,
                             dfront_f, nx, ny, nz, rho, mx, my, mz, cn2, ci2, smax_out, stream);
}

extern "C" int nv_ionize_apply_h(const float* rfronts, int nu, int nphi, const float* star_xyz, int nstars, float dx,
                                 float width, float* xi, uint16_t* dfront_h, float* dfront_f, int nx, int ny, int nz,
                                 const float* rho, const uint16_t* vx, const uint16_t* vy, const uint16_t* vz,
                                 float cn2, float ci2, float* smax_out, void* stream) {
  return ionize_apply_launch("nv_ionize_apply_h", NV_ION_VEL, rfronts, nu, nphi, star_xyz, nstars, dx, width, xi, dfront_h,
                             dfront_f, nx, ny, nz, rho, vx, vy, vz, cn2, ci2, smax_out, stream);
}

extern "C" int nv_ionize_apply_ex(const NvIonDesc* d, int form, void* stream) {
  clear_error();
  NV_REQUIRE(d, "nv_ionize_apply_ex: null descriptor");
  return ionize_apply_launch("nv_ionize_apply_ex", form, d->rfronts, d->nu, d->nphi, d->star_xyz, d->nstars, d->dx,
                             d->width, d->xi, d->dfront_h, d->dfront_f, d->nx, d->ny, d->nz, d->state[0], d->state[1],
                             d->state[2], d->state[3], d->cn2, d->ci2, d->smax_out, stream);
}

extern "C" int nv_cast_fuv(const float* rho, const float* xi, int nx, int ny, int nz, float sx, float sy, float sz,
                           float kappa, float dr, int nsteps, float dx, float binr, float bx, float by, float bz,
                           float* att, int nr, int nu, int nphi, void* stream) {
  clear_error();
  NV_REQUIRE(rho && xi && att, "nv_cast_fuv: null pointer");
  NV_REQUIRE(nx > 0 && ny > 0 && nz > 0 && nr > 0 && nu > 0 && nphi > 0, "nv_cast_fuv: bad sizes");
  NV_REQUIRE(dr > 0.0f && dx > 0.0f && binr > 0.0f, "nv_cast_fuv: dr, dx, binr must be > 0");
  NV_REQUIRE((long long)nx * ny * nz < (1LL << 32), "nv_cast_fuv: grids above 2^32 cells are not supported");
  const long long n = (long long)nu * nphi;
  dispatch_sampler(nz, (long long)nx * ny * nz, (long long)nx * ny, dx, [&](auto pair, auto off, auto sc) {
    constexpr bool P = decltype(pair)::value;
    hipLaunchKernelGGL((cast_fuv_kernel<P, decltype(off), decltype(sc)>), dim3(blocks_for(n, 256)), dim3(256), 0,
                       as_stream(stream), rho, xi, nx, ny, nz, sx, sy, sz, kappa, dr, nsteps, sc, binr, bx, by, bz,
                       att, nr, nu, nphi);
  });
  return check_launch("nv_cast_fuv");
}

extern "C" int nv_apply_fuv(const float* att, int nr, int nu, int nphi, float sx, float sy, float sz, float lum,
                            float dx, float binr, float* g0, int nx, int ny, int nz, void* stream) {
  clear_error();
  NV_REQUIRE(att && g0, "nv_apply_fuv: null pointer");
  NV_REQUIRE(nx > 0 && ny > 0 && nz > 0 && nr > 0 && nu > 0 && nphi > 0, "nv_apply_fuv: bad sizes");
  NV_REQUIRE(binr >= 1.17549435e-38f && binr <= 3.40282347e+38f, "nv_apply_fuv: binr must be a positive normal float");
  const long long n = (long long)nx * ny * nz;
  hipLaunchKernelGGL(apply_fuv_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, as_stream(stream), att, nr, nu, nphi,
                     sx, sy, sz, lum, dx, Recip{1.0 / (double)binr, binr}, g0, nx, ny, nz);
  return check_launch("nv_apply_fuv");
}

extern "C" int nv_fuv_apply(const float* atts, int nr, int nu, int nphi, const float* star_xyz, const float* lums,
                            int nstars, float dx, float binr, uint16_t* g0_h, float* g0_f, int nx, int ny, int nz,
                            void* stream) {
  clear_error();
  NV_REQUIRE(nstars == 0 || (atts && star_xyz && lums), "nv_fuv_apply: null pointer");
  NV_REQUIRE((g0_h != nullptr) != (g0_f != nullptr), "nv_fuv_apply: pass exactly one g0 buffer");
  NV_REQUIRE(nx > 0 && ny > 0 && nz > 0 && nr > 0 && nu > 0 && nphi > 0 && nstars >= 0, "nv_fuv_apply: bad sizes");
  NV_REQUIRE(binr >= 1.17549435e-38f && binr <= 3.40282347e+38f, "nv_fuv_apply: binr must be a positive normal float");
  const long long n = (long long)nx * ny * nz;
  hipLaunchKernelGGL(fuv_apply_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, as_stream(stream), atts, nr, nu, nphi,
                     star_xyz, lums, nstars, dx, Recip{1.0 / (double)binr, binr}, g0_h, g0_f, nx, ny, nz);
  return check_launch("nv_fuv_apply");
}

// ---------------------------------------------------------------------------------------- batch plans
// Plan layout (host blob, uploaded once): a 64-byte header, the jobs' kernel arguments, then start[0..n]; a ray
// plan continues with its brick jobs (at most one per job) and their block prefix bstart[0..nb].
namespace nv {
// Each job's front range from its workgroups' partial ranges: one workgroup per job.
__global__ __launch_bounds__(256) void rf_range_kernel(const RayJob* __restrict__ jobs) {
  __shared__ float lo_[4], hi_[4];
  const RayJob& J = jobs[blockIdx.x];
  if (!J.rf_range) return;
  const int nb = (int)ray_blocks(J.nu, J.nphi);
  float lo = __builtin_nanf(""), hi = lo;
  for (int b = threadIdx.x; b < nb; b += 256) {
    lo = fminf(lo, J.rf_range[2 + 2 * b]);
    hi = fmaxf(hi, J.rf_range[3 + 2 * b]);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    lo = fminf(lo, __shfl_xor(lo, off, 64));
    hi = fmaxf(hi, __shfl_xor(hi, off, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    lo_[threadIdx.x >> 6] = lo;
    hi_[threadIdx.x >> 6] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    J.rf_range[0] = fminf(fminf(lo_[0], lo_[1]), fminf(lo_[2], lo_[3]));
    J.rf_range[1] = fmaxf(fmaxf(hi_[0], hi_[1]), fmaxf(hi_[2], hi_[3]));
  }
}

struct RayBatchHeader {
  uint32_t magic;
  int n, total, pair, off32;
  int nb, btotal;          // ray plans: brick jobs and their blocks
  int ranges;              // some job asks for its front range (ray plans) or member carries them (ionize plans)
  int fast;                // ionize plans: the fast arithmetic (NV_ION_FAST)
  int xcd_run;             // ionize plans: the longest XCD run of tiles (the launch has 8 xcd_run workgroups)
  int pad[6];
};
static_assert(sizeof(RayBatchHeader) == 64, "plan header layout");
constexpr uint32_t kRayMagic = 0x4e565231u;   // "NVR1"
constexpr uint32_t kIonMagic = 0x4e564931u;   // "NVI1"

template <class T>
static long long batch_bytes(int n) {
  return (long long)sizeof(RayBatchHeader) + (long long)n * (long long)sizeof(T) + 4LL * (n + 1);
}
}  // namespace nv

static long long ray_plan_bytes(int n) {
  return batch_bytes<RayJob>(n) + (long long)n * (long long)sizeof(BrickJob) + 4LL * (n + 1);
}

extern "C" long long nv_cast_rays_batch_plan_size(int n) { return n > 0 ? ray_plan_bytes(n) : -1; }

extern "C" int nv_cast_rays_batch_plan(const NvRayDesc* jobs, int n, void* plan, long long plan_bytes) {
  clear_error();
  NV_REQUIRE(jobs && plan && n > 0, "nv_cast_rays_batch_plan: need n > 0 jobs and a plan buffer");
  NV_REQUIRE(plan_bytes >= ray_plan_bytes(n), "nv_cast_rays_batch_plan: plan needs %lld bytes", ray_plan_bytes(n));
  RayBatchHeader H = {};
  RayJob* out = reinterpret_cast<RayJob*>(static_cast<char*>(plan) + sizeof(H));
  int* start = reinterpret_cast<int*>(out + n);
  BrickJob* bj = reinterpret_cast<BrickJob*>(start + n + 1);
  int* bstart = reinterpret_cast<int*>(bj + n);
  long long total = 0, btotal = 0;
  int nb = 0;
  bool pair = true, off32 = true, ranges = false;
  for (int s = 0; s < n; ++s) {
    const NvRayDesc& d = jobs[s];
    NV_REQUIRE(d.rho && d.rfront, "nv_cast_rays_batch_plan: job %d: null pointer", s);
    NV_REQUIRE(d.nx > 0 && d.ny > 0 && d.nz > 0 && d.nu > 0 && d.nphi > 0, "nv_cast_rays_batch_plan: job %d: bad sizes",
               s);
    NV_REQUIRE(d.dr > 0.0f && d.dx > 0.0f, "nv_cast_rays_batch_plan: job %d: dr and dx must be > 0", s);
    const long long cells = (long long)d.nx * d.ny * d.nz;
    NV_REQUIRE(cells < (1LL << 32), "nv_cast_rays_batch_plan: job %d: grids above 2^32 cells are not supported", s);
    pair = pair && d.nz >= 2;
    // 32-bit byte offsets: 4-byte elements up to 2^30 cells, 16-byte bricks up to 2^28
    if (d.bricks) {   // the padded brick grid: 32-bit offsets up to 2^28 bricks, 24-bit products below 2^24 rows
      const long long prow = (long long)(d.nx + 2) * (d.ny + 1);
      off32 = off32 && prow * (d.nz + 1) <= (1LL << 28) && prow < (1LL << 24);
    } else {
      off32 = off32 && cells <= (1LL << 30) && (long long)d.nx * d.ny < (1LL << 24);
    }
    RayJob J;
    fill_ray_job(d.rho, d.nx, d.ny, d.nz, d.sx, d.sy, d.sz, d.Q, d.dr, d.nsteps, d.dx, d.rcap, d.bx, d.by, d.bz,
                 d.rfront, d.nu, d.nphi, d.wide != 0, J);
    NV_REQUIRE(!d.dirs || ((uintptr_t)d.dirs & 15) == 0, "nv_cast_rays_batch_plan: job %d: dirs must be 16-byte aligned",
               s);
    J.dirs = reinterpret_cast<const float4*>(d.dirs);
    J.rf_range = d.rf_range;
    ranges = ranges || d.rf_range != nullptr;
    if (d.bricks) {
      const char* r = reinterpret_cast<const char*>(d.rho);
      const char* b = reinterpret_cast<const char*>(d.bricks);
      NV_REQUIRE(((uintptr_t)d.bricks & 15) == 0, "nv_cast_rays_batch_plan: job %d: bricks must be 16-byte aligned", s);
      NV_REQUIRE(b + nv_ray_bricks_size(d.nx, d.ny, d.nz) <= r || r + 4 * cells <= b,
                 "nv_cast_rays_batch_plan: job %d: bricks overlap rho", s);
      NV_REQUIRE(d.ny <= 65534 && d.nx <= 65533 && nv_ray_bricks_size(d.nx, d.ny, d.nz) / 16 < (1LL << 32),
                 "nv_cast_rays_batch_plan: job %d: bad sizes", s);
      J.bricks = reinterpret_cast<const float4*>(d.bricks);
      // one brick job per distinct brick array (the jobs of one grid's stars share it)
      bool seen = false;
      for (int q = 0; q < nb; ++q) {
        if (bj[q].B == reinterpret_cast<float4*>(d.bricks)) {
          NV_REQUIRE(bj[q].f == d.rho && bj[q].nx == d.nx && bj[q].ny == d.ny && bj[q].nz == d.nz,
                     "nv_cast_rays_batch_plan: job %d: one brick array for two different grids", s);
          seen = true;
        }
      }
      if (!seen) {
        BrickJob B;
        B.f = d.rho;
        B.B = reinterpret_cast<float4*>(d.bricks);
        B.nx = d.nx; B.ny = d.ny; B.nz = d.nz;
        B.pad = 0;
        memcpy(&bj[nb], &B, sizeof(B));
        const long long nbricks = (long long)(d.nx + 2) * (d.ny + 1) * (d.nz + 1);
        NV_REQUIRE(nbricks < (1LL << 31), "nv_cast_rays_batch_plan: job %d: above 2^31 bricks", s);
        bstart[nb++] = (int)btotal;
        btotal += (nbricks + 255) / 256;
        NV_REQUIRE(btotal < (1LL << 31) - 1, "nv_cast_rays_batch_plan: too many brick blocks");
      }
    }
    memcpy(&out[s], &J, sizeof(J));
    start[s] = (int)total;
    total += ray_blocks(d.nu, d.nphi);
    NV_REQUIRE(total < (1LL << 31) - 1, "nv_cast_rays_batch_plan: too many rays");
  }
  start[n] = (int)total;
  bstart[nb] = (int)btotal;
  H.magic = kRayMagic;
  H.n = n;
  H.total = (int)total;
  H.pair = pair ? 1 : 0;
  H.off32 = off32 ? 1 : 0;
  H.nb = nb;
  H.btotal = (int)btotal;
  H.ranges = ranges ? 1 : 0;
  memcpy(plan, &H, sizeof(H));
  return NV_OK;
}

extern "C" int nv_cast_rays_batch(const void* plan_host, const void* plan_dev, void* stream) {
  clear_error();
  NV_REQUIRE(plan_host && plan_dev, "nv_cast_rays_batch: null plan");
  RayBatchHeader H;
  memcpy(&H, plan_host, sizeof(H));
  NV_REQUIRE(H.magic == kRayMagic && H.n > 0, "nv_cast_rays_batch: not a ray batch plan");
  const RayJob* jobs = reinterpret_cast<const RayJob*>(static_cast<const char*>(plan_dev) + sizeof(H));
  const int* start = reinterpret_cast<const int*>(jobs + H.n);
  if (H.nb > 0) {
    const BrickJob* bj = reinterpret_cast<const BrickJob*>(start + H.n + 1);
    const int* bstart = reinterpret_cast<const int*>(bj + H.n);
    hipLaunchKernelGGL(ray_bricks_batch_kernel, dim3(H.btotal), dim3(256), 0, as_stream(stream), bj, bstart, H.nb,
                       H.btotal);
  }
  auto pick = [&](auto brick) {
    constexpr bool BR = decltype(brick)::value;
    return H.pair ? (H.off32 ? cast_rays_batch_kernel<true, uint32_t, BR> : cast_rays_batch_kernel<true, uint64_t, BR>)
                  : (H.off32 ? cast_rays_batch_kernel<false, uint32_t, BR> : cast_rays_batch_kernel<false, uint64_t, BR>);
  };
  auto kern = H.nb > 0 ? pick(std::true_type{}) : pick(std::false_type{});
  hipLaunchKernelGGL(kern, dim3(H.total), dim3(256), 0, as_stream(stream), jobs, start, H.n, H.total);
  if (H.ranges) hipLaunchKernelGGL(rf_range_kernel, dim3(H.n), dim3(256), 0, as_stream(stream), jobs);
  return check_launch("nv_cast_rays_batch");
}

// an ionize plan: the header, the members' arguments, start[0..n], then the XCD cuts xoff[0..8]
static long long ion_plan_bytes(int n) { return batch_bytes<IonArgs>(n) + 4LL * 9; }

extern "C" long long nv_ionize_apply_batch_plan_size(int n) { return n > 0 ? ion_plan_bytes(n) : -1; }

// The relative work of one tile of a member with k stars, for cutting the tile list into runs of equal work: the
// state loads, stores and smoothing are the same for any k; each further star costs a full evaluation per cell
// without culling, and with it only the distance tests of the cells it provably leaves unchanged (the 64-member
// batch: about 1.03 full evaluations per cell).
static double tile_work(int nstars, bool culls) {
  if (nstars <= 1) return 1.0;
  return 1.0 + (culls ? 0.25 : 0.8) * (double)(nstars - 1);
}

extern "C" int nv_ionize_apply_batch_plan(const NvIonDesc* members, int n, int form, void* plan,
                                          long long plan_bytes) {
  clear_error();
  NV_REQUIRE(form >= 0 && form <= (NV_ION_VEL | NV_ION_FAST), "nv_ionize_apply_batch_plan: bad form %d", form);
  NV_REQUIRE(members && plan && n > 0, "nv_ionize_apply_batch_plan: need n > 0 members and a plan buffer");
  NV_REQUIRE(plan_bytes >= ion_plan_bytes(n), "nv_ionize_apply_batch_plan: plan needs %lld bytes", ion_plan_bytes(n));
  long long jk_total = 0;
  for (int s = 0; s < n; ++s) {
    NV_REQUIRE(members[s].ny > 0 && members[s].nz > 0, "nv_ionize_apply_batch_plan: member %d: bad grid", s);
    jk_total += ion_jk(members[s].ny, members[s].nz);
  }
  RayBatchHeader H = {};
  IonArgs* out = reinterpret_cast<IonArgs*>(static_cast<char*>(plan) + sizeof(H));
  int* start = reinterpret_cast<int*>(out + n);
  int* xoff = start + n + 1;
  const bool sm = members[0].smax_out != nullptr;
  bool ranges = false;
  long long total = 0;
  double work = 0.0;
  char name[64];
  for (int s = 0; s < n; ++s) {
    const NvIonDesc& d = members[s];
    snprintf(name, sizeof(name), "nv_ionize_apply_batch_plan: member %d", s);
    NV_REQUIRE((d.smax_out != nullptr) == sm, "%s: smax_out must be set for all members or none", name);
    IonArgs A;
    const int rc = fill_ionize(name, d.rfronts, d.nu, d.nphi, d.star_xyz, d.nstars, d.dx, d.width, d.xi, d.dfront_h,
                               d.dfront_f, d.nx, d.ny, d.nz, d.state[0], d.state[1], d.state[2], d.state[3], d.cn2,
                               d.ci2, d.smax_out, jk_total, kIonBatchDepth, A);
    if (rc != NV_OK) return rc;
    A.rf_ranges = d.rf_ranges;
    ranges = ranges || d.rf_ranges != nullptr;
    memcpy(&out[s], &A, sizeof(A));
    start[s] = (int)total;
    total += A.tiles;
    work += (double)A.tiles * tile_work(d.nstars, d.rf_ranges != nullptr);
    NV_REQUIRE(total < (1LL << 30), "nv_ionize_apply_batch_plan: more than 2^30 tiles");
  }
  start[n] = (int)total;
  // XCD x takes the tiles from the first one at which the work done reaches x / 8 of the whole, to the next cut
  xoff[0] = 0;
  int s = 0;
  double before = 0.0;   // the work of members 0 .. s-1
  for (int x = 1; x < 8; ++x) {
    const double target = work * x / 8.0;
    for (;;) {
      const double w = tile_work(members[s].nstars, members[s].rf_ranges != nullptr);
      const double mw = (double)(start[s + 1] - start[s]) * w;
      if (s + 1 < n && before + mw <= target) {
        before += mw;
        ++s;
        continue;
      }
      const long long t = start[s] + (long long)std::ceil((target - before) / w);
      xoff[x] = (int)std::max<long long>(xoff[x - 1], std::min<long long>(t, start[s + 1]));
      break;
    }
  }
  xoff[8] = (int)total;
  int run = 0;
  for (int x = 0; x < 8; ++x) {
    NV_REQUIRE(xoff[x] <= xoff[x + 1], "nv_ionize_apply_batch_plan: XCD cuts out of order");
    run = std::max(run, xoff[x + 1] - xoff[x]);
  }
  H.magic = kIonMagic;
  H.n = n;
  H.total = (int)total;
  H.pair = sm ? 1 : 0;     // reused: the CFL reduction is on
  H.off32 = (form & NV_ION_VEL) ? 1 : 0;   // reused: binary16-velocity state
  H.fast = (form & NV_ION_FAST) ? 1 : 0;
  H.ranges = ranges ? 1 : 0;
  H.xcd_run = run;
  memcpy(plan, &H, sizeof(H));
  return NV_OK;
}

extern "C" int nv_ionize_apply_batch(const void* plan_host, const void* plan_dev, void* stream) {
  clear_error();
  NV_REQUIRE(plan_host && plan_dev, "nv_ionize_apply_batch: null plan");
  RayBatchHeader H;
  memcpy(&H, plan_host, sizeof(H));
  NV_REQUIRE(H.magic == kIonMagic && H.n > 0, "nv_ionize_apply_batch: not an ionize batch plan");
  const IonArgs* args = reinterpret_cast<const IonArgs*>(static_cast<const char*>(plan_dev) + sizeof(H));
  const int* start = reinterpret_cast<const int*>(args + H.n);
  const int* xoff = start + H.n + 1;
  auto pick = [&](auto cull, auto fast) {
    constexpr bool C = decltype(cull)::value, F = decltype(fast)::value;
    return !H.pair ? ionize_apply_batch_kernel<false, false, C, F>
                   : (H.off32 ? ionize_apply_batch_kernel<true, true, C, F>
                              : ionize_apply_batch_kernel<true, false, C, F>);
  };
  auto by_cull = [&](auto fast) { return H.ranges ? pick(std::true_type{}, fast) : pick(std::false_type{}, fast); };
  auto kern = H.fast ? by_cull(std::true_type{}) : by_cull(std::false_type{});
  NV_REQUIRE(H.xcd_run > 0, "nv_ionize_apply_batch: the plan has no tiles");
  hipLaunchKernelGGL(kern, dim3(8 * H.xcd_run), dim3(IK, IJ), 0, as_stream(stream), args, start, H.n, xoff);
  return check_launch("nv_ionize_apply_batch");
}
