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
#include "nv_common.h"
#include "nv_math.h"
#include "nv_divide.h"

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

// Two z-neighbours (i, i+1) in one 8-byte load (4-byte aligned, which global loads allow).
__device__ __forceinline__ float2 ld_pair(const float* __restrict__ f, unsigned i) {
  typedef float f2u __attribute__((ext_vector_type(2), aligned(4)));
  const f2u v = *reinterpret_cast<const f2u*>(f + i);
  return make_float2(v.x, v.y);
}

// How a sampler turns a world coordinate into cell units: p / dx through the shared double reciprocal (the
// identical correctly rounded quotient, nv_divide.h), or p * (1/dx) when dx is a power of two, which rounds the
// same real number and so gives the same float.
struct CellScale {
  double rdx;
  float inv;   // 1/dx if dx = 2^e, else 0
  __device__ __forceinline__ float operator()(float p) const { return inv != 0.0f ? p * inv : div_r(p, rdx); }
};

// trilinear sample at world p, 0 outside [-1/2, n-1/2] in cell units (sample_n, nebula3d.py:195-226), with
// 32-bit cell indices (grids up to 2^32 cells). PAIR (nz >= 2): the corners at ka = clamp(k0) and
// kb = clamp(k0+1) come from one load of (kp, kp+1), kp = clamp(k0, 0, nz-2): k0 = -1 gives ka = kb = 0 (the
// pair's first value), k0 = nz-1 gives ka = kb = nz-1 (its second), so every corner is the value the separate
// clamped loads read.
template <bool PAIR>
__device__ __forceinline__ float sample_r(const float* __restrict__ f, int nx, int ny, int nz, float px, float py,
                                          float pz, const CellScale& cs) {
  const float gx = cs(px) - 0.5f, gy = cs(py) - 0.5f, gz = cs(pz) - 0.5f;
  if (gx < -0.5f || gy < -0.5f || gz < -0.5f) return 0.0f;
  if (gx > (float)nx - 0.5f || gy > (float)ny - 0.5f || gz > (float)nz - 0.5f) return 0.0f;
  const int i0 = (int)floorf(gx), j0 = (int)floorf(gy), k0 = (int)floorf(gz);
  const float fx = gx - (float)i0, fy = gy - (float)j0, fz = gz - (float)k0;
  const unsigned ia = clampi(i0, 0, nx - 1), ib = clampi(i0 + 1, 0, nx - 1);
  const unsigned ja = clampi(j0, 0, ny - 1), jb = clampi(j0 + 1, 0, ny - 1);
  const unsigned r00 = (ia * ny + ja) * nz, r10 = (ib * ny + ja) * nz;
  const unsigned r01 = (ia * ny + jb) * nz, r11 = (ib * ny + jb) * nz;
  float c000, c100, c010, c110, c001, c101, c011, c111;
  if constexpr (PAIR) {
    const unsigned kp = clampi(k0, 0, nz - 2);
    const bool a_second = k0 > nz - 2, b_first = k0 < 0;
    const float2 p00 = ld_pair(f, r00 + kp), p10 = ld_pair(f, r10 + kp);
    const float2 p01 = ld_pair(f, r01 + kp), p11 = ld_pair(f, r11 + kp);
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
    c000 = f[r00 + ka]; c100 = f[r10 + ka]; c010 = f[r01 + ka]; c110 = f[r11 + ka];
    c001 = f[r00 + kb]; c101 = f[r10 + kb]; c011 = f[r01 + kb]; c111 = f[r11 + kb];
  }
  const float c00 = c000 + (c100 - c000) * fx;
  const float c10 = c010 + (c110 - c010) * fx;
  const float c01 = c001 + (c101 - c001) * fx;
  const float c11 = c011 + (c111 - c011) * fx;
  const float c0 = c00 + (c10 - c00) * fy;
  const float c1 = c01 + (c11 - c01) * fy;
  return c0 + (c1 - c0) * fz;
}

// Direction cell of thread t: each wavefront takes an 8 x 8 patch of (u, phi) cells, so its 64 rays stay
// within a few cells of each other and their corner gathers share cache lines.
__device__ __forceinline__ bool ray_of_thread(long long t, int nu, int nphi, int* iu, int* ip) {
  const int tp = (nphi + 7) >> 3;
  const long long patch = t >> 6;
  const int lane = (int)(t & 63);
  *iu = (int)(patch / tp) * 8 + (lane >> 3);
  *ip = (int)(patch % tp) * 8 + (lane & 7);
  return *iu < nu && *ip < nphi;
}

// cast_rays (nebula3d.py:247-270). Two samples per iteration: the corner loads of sample n+1 are in flight
// while sample n is tested, and the column is still accumulated sample by sample in the reference order.
// WIDE: a ray whose marched length (n1 - n0) dr exceeds 512 cells keeps its column in FP64 (the test
// col + inc > Q and the interpolation of rf in double); shorter rays are unchanged.
template <bool PAIR, bool WIDE>
__global__ __launch_bounds__(256) void cast_rays_kernel(const float* __restrict__ rho, int nx, int ny, int nz,
                                                        float sx, float sy, float sz, float Q, float dr,
                                                        int nsteps, CellScale cs, float dx, float rcap, float bx,
                                                        float by, float bz, float* __restrict__ rfront, int nu,
                                                        int nphi) {
  int iu, ip;
  if (!ray_of_thread((long long)blockIdx.x * blockDim.x + threadIdx.x, nu, nphi, &iu, &ip)) return;
  float d[3];
  ray_dir(iu, ip, nu, nphi, d);
  const float st[3] = {sx, sy, sz}, bm[3] = {bx, by, bz};
  float t0, t1;
  box_range(st, d, bm, &t0, &t1);
  const int n0 = (int)(t0 / dr);
  const int n1 = min(nsteps, (int)(t1 / dr) + 2);
  const bool wide = WIDE && (float)(n1 - n0) * dr > 512.0f * dx;
  float col = 0.0f, rf = rcap;
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
    const float na = sample_r<PAIR>(rho, nx, ny, nz, sx + d[0] * ra, sy + d[1] * ra, sz + d[2] * ra, cs);
    const float nb = sample_r<PAIR>(rho, nx, ny, nz, sx + d[0] * rb, sy + d[1] * rb, sz + d[2] * rb, cs);
    if (step(ra, na * na * ra * ra * dr)) break;
    if (n + 1 >= n1) break;
    if (step(rb, nb * nb * rb * rb * dr)) break;
  }
  rfront[(long long)iu * nphi + ip] = rf;
}

// cell centre seen from a star: r, u = cos(theta), phi in [0, 2pi) (nebula3d.py:284-290)
__device__ __forceinline__ void cell_sph(int i, int j, int k, float sx, float sy, float sz, float dx, float* r_o,
                                         float* u_o, float* phi_o) {
  const float r0 = ((float)i + 0.5f) * dx - sx, r1 = ((float)j + 0.5f) * dx - sy, r2 = ((float)k + 0.5f) * dx - sz;
  const float r = fmaxf(sqrtf((r0 * r0 + r1 * r1) + r2 * r2), 1e-6f);
  float phi = nvm::atan2f_det(r1, r0);
  if (phi < 0.0f) phi = phi + 2.0f * PI_F;
  *r_o = r;
  *u_o = r2 / r;
  *phi_o = phi;
}

// bilinear front radius at direction (u, phi) (nebula3d.py:291-301, rf_at 273-278)
__device__ __forceinline__ float front_at(const float* __restrict__ rfront, int nu, int nphi, float u, float phi) {
  const float fu = (u + 1.0f) * 0.5f * (float)nu - 0.5f;
  const float fp = phi / (2.0f * PI_F) * (float)nphi - 0.5f;
  const int iu = (int)floorf(fu), ip = (int)floorf(fp);
  const float tu = fu - (float)iu, tp = fp - (float)ip;
  const int ua = clampi(iu, 0, nu - 1), ub = clampi(iu + 1, 0, nu - 1);
  const int pa = ((ip % nphi) + nphi) % nphi, pb = (((ip + 1) % nphi) + nphi) % nphi;
  const float* ra = rfront + (long long)ua * nphi;
  const float* rb = rfront + (long long)ub * nphi;
  const float av = ra[pa] * (1.0f - tp) + ra[pb] * tp;
  const float bv = rb[pa] * (1.0f - tp) + rb[pb] * tp;
  return av * (1.0f - tu) + bv * tu;
}

// apply_front (nebula3d.py:281-307)
__global__ __launch_bounds__(256) void apply_front_kernel(const float* __restrict__ rfront, int nu, int nphi,
                                                          float sx, float sy, float sz, float dx, float width,
                                                          float* xi, float* dfront, int nx, int ny, int nz) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long long)nx * ny * nz) return;
  const int k = (int)(q % nz);
  const long long ij = q / nz;
  const int j = (int)(ij % ny), i = (int)(ij / ny);
  float r, u, phi;
  cell_sph(i, j, k, sx, sy, sz, dx, &r, &u, &phi);
  const float rf = front_at(rfront, nu, nphi, u, phi);
  const float val = fminf(fmaxf((rf - r) / width + 0.5f, 0.0f), 1.0f);
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
  const float m = (src[cell(clampi(i - 1, 0, nx - 1), j, k, ny, nz)] + src[cell(clampi(i + 1, 0, nx - 1), j, k, ny, nz)] +
                   src[cell(i, clampi(j - 1, 0, ny - 1), k, ny, nz)] + src[cell(i, clampi(j + 1, 0, ny - 1), k, ny, nz)] +
                   src[cell(i, j, clampi(k - 1, 0, nz - 1), ny, nz)] + src[cell(i, j, clampi(k + 1, 0, nz - 1), ny, nz)]) /
                  6.0f;
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
constexpr int IK = 64, IJ = 8, INT_ = IK * IJ;     // tile: 64 z x 8 y, marching along x
constexpr int XK = IK + 2, XJ = IJ + 2, XN = XK * XJ;
constexpr int XR = 4;                               // ring of unsmoothed-xi planes (i-1, i, i+1 read; i+2 written)

struct IonArgs {
  const float* rfronts;
  const float* star_xyz;
  float* xi;
  uint16_t* df_h;
  float* df_f;
  const void *f0, *f1, *f2, *f3;   // state for the CFL max: (rho, m) float32, or (rho, v) binary16
  float* smax;
  int nu, nphi, nstars, nx, ny, nz, ci;
  float dx, width, cn2, ci2;
};

template <bool SMAX, bool VEL>
__global__ __launch_bounds__(INT_) void ionize_apply_kernel(IonArgs A) {
  __shared__ float sx_[XR][XN];   // unsmoothed xi (max over stars of the front ramp)
  __shared__ float sd_[XR][XN];   // fp32 dfront (min over stars of r - rf)
  __shared__ float red[INT_ / 64];
  const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * IK + tx;
  const int nx = A.nx, ny = A.ny, nz = A.nz;
  const int k0 = blockIdx.x * IK, j0 = blockIdx.y * IJ, i0 = blockIdx.z * A.ci;
  const int i1 = min(i0 + A.ci, nx);
  const int k = k0 + tx, j = j0 + ty;
  const bool active = k < nz && j < ny;
  const long long tab = (long long)A.nu * A.nphi;

  // unsmoothed xi and fp32 dfront of plane p (clamped cell positions, as ldc in smooth_xi)
  auto make_plane = [&](int p) {
    const int s = ((p % XR) + XR) % XR;
    const int ic = clampi(p, 0, nx - 1);
    for (int t = tid; t < XN; t += INT_) {
      const int jj = t / XK, kk = t - jj * XK;
      const int jc = clampi(j0 + jj - 1, 0, ny - 1), kc = clampi(k0 + kk - 1, 0, nz - 1);
      float x = 0.0f, df = 1e3f;   // xi.zero_(), df.fill_(1e3) (sim.py:930-931)
      for (int st = 0; st < A.nstars; ++st) {
        const float ssx = A.star_xyz[3 * st], ssy = A.star_xyz[3 * st + 1], ssz = A.star_xyz[3 * st + 2];
        float r, u, phi;
        cell_sph(ic, jc, kc, ssx, ssy, ssz, A.dx, &r, &u, &phi);
        const float rf = front_at(A.rfronts + st * tab, A.nu, A.nphi, u, phi);
        const float val = fminf(fmaxf((rf - r) / A.width + 0.5f, 0.0f), 1.0f);
        if (val > x) x = val;
        df = fminf(df, r - rf);
      }
      sx_[s][t] = x;
      sd_[s][t] = df;
    }
  };

  make_plane(i0 - 1);
  make_plane(i0);
  float tmax = 0.0f;
  for (int i = i0; i < i1; ++i) {
    make_plane(i + 1);
    __syncthreads();
    if (active) {
      const int c = (ty + 1) * XK + tx + 1;
      const float* pc = sx_[((i % XR) + XR) % XR];
      const float* pm = sx_[(((i - 1) % XR) + XR) % XR];
      const float* pp = sx_[(((i + 1) % XR) + XR) % XR];
      // neighbours in the order of nebula3d.py:313-314
      const float m = (pm[c] + pp[c] + pc[c - XK] + pc[c + XK] + pc[c - 1] + pc[c + 1]) / 6.0f;
      const float xs = 0.5f * pc[c] + 0.5f * m;
      const long long o = cell(i, j, k, ny, nz);
      A.xi[o] = xs;
      const float df = sd_[((i % XR) + XR) % XR][c];
      if (A.df_h) A.df_h[o] = f32_to_f16_bits(df);
      else A.df_f[o] = df;
      if constexpr (SMAX) {
        float s;
        if constexpr (VEL) {
          s = ((fabsf(ldf((const uint16_t*)A.f1, o)) + fabsf(ldf((const uint16_t*)A.f2, o))) +
               fabsf(ldf((const uint16_t*)A.f3, o))) + sqrtf(A.cn2 + xs * (A.ci2 - A.cn2));
        } else {
          const float rr = ((const float*)A.f0)[o];
          s = ((fabsf(((const float*)A.f1)[o] / rr) + fabsf(((const float*)A.f2)[o] / rr)) +
               fabsf(((const float*)A.f3)[o] / rr)) + sqrtf(A.cn2 + xs * (A.ci2 - A.cn2));
        }
        tmax = fmaxf(tmax, s);
      }
    }
  }
  if constexpr (SMAX) block_max_atomic(tmax, red, A.smax);
}

// ---------------------------------------------------------------------------------------- FUV
// cast_fuv (nebula3d.py:320-352)
template <bool PAIR>
__global__ __launch_bounds__(256) void cast_fuv_kernel(const float* __restrict__ rho, const float* __restrict__ xi,
                                                       int nx, int ny, int nz, float sx, float sy, float sz,
                                                       float kappa, float dr, int nsteps, CellScale cs, float binr,
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
    const float n_tot = sample_r<PAIR>(rho, nx, ny, nz, qx, qy, qz, cs);
    const float x_i = sample_r<PAIR>(xi, nx, ny, nz, qx, qy, qz, cs);
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

// trilinear att at (r/binr - 1/2, u, phi) divided by max(r^2, 4e-4) (nebula3d.py:362-387)
__device__ __forceinline__ float fuv_at(const float* __restrict__ att, int nr, int nu, int nphi, float r, float u,
                                        float phi, float binr) {
  const float fr = r / binr - 0.5f;
  const float fu = (u + 1.0f) * 0.5f * (float)nu - 0.5f;
  const float fp = phi / (2.0f * PI_F) * (float)nphi - 0.5f;
  const int ir = (int)floorf(fr), iu = (int)floorf(fu), ip = (int)floorf(fp);
  const float tr = fr - (float)ir, tu = fu - (float)iu, tp = fp - (float)ip;
  float acc = 0.0f;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const float w = (a == 0 ? 1.0f - tr : tr) * (b == 0 ? 1.0f - tu : tu) * (c == 0 ? 1.0f - tp : tp);
        const int rr = clampi(ir + a, 0, nr - 1), uu = clampi(iu + b, 0, nu - 1);
        const int pp = (((ip + c) % nphi) + nphi) % nphi;
        acc = acc + w * att[((long long)rr * nu + uu) * nphi + pp];
      }
  return acc;
}

// apply_fuv (nebula3d.py:355-388)
__global__ __launch_bounds__(256) void apply_fuv_kernel(const float* __restrict__ att, int nr, int nu, int nphi,
                                                        float sx, float sy, float sz, float lum, float dx,
                                                        float binr, float* g0, int nx, int ny, int nz) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long long)nx * ny * nz) return;
  const int k = (int)(q % nz);
  const long long ij = q / nz;
  const int j = (int)(ij % ny), i = (int)(ij / ny);
  float r, u, phi;
  cell_sph(i, j, k, sx, sy, sz, dx, &r, &u, &phi);
  const float acc = fuv_at(att, nr, nu, nphi, r, u, phi, binr);
  const float r2 = fmaxf(r * r, 0.0004f);
  g0[q] = g0[q] + lum * acc / r2;
}

// the per-cell part of Sim.fuv_field (sim.py:955-978), all stars in order, one store
__global__ __launch_bounds__(256) void fuv_apply_kernel(const float* __restrict__ atts, int nr, int nu, int nphi,
                                                        const float* __restrict__ star_xyz,
                                                        const float* __restrict__ lums, int nstars, float dx,
                                                        float binr, uint16_t* g0_h, float* g0_f, int nx, int ny,
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
    const float acc = fuv_at(atts + s * tab, nr, nu, nphi, r, u, phi, binr);
    const float r2 = fmaxf(r * r, 0.0004f);
    g = g + lums[s] * acc / r2;
  }
  if (g0_h) g0_h[q] = f32_to_f16_bits(g);
  else g0_f[q] = g;
}

inline int blocks_for(long long n, int bs) { return (int)((n + bs - 1) / bs); }

}  // namespace nv

using namespace nv;

static CellScale cell_scale(float dx) {
  CellScale cs;
  cs.rdx = 1.0 / (double)dx;   // correctly rounded: |rdx dx - 1| <= 2^-53, inside the bound div_r needs
  cs.inv = pow2_reciprocal(dx);
  return cs;
}

static int cast_rays_launch(const char* name, bool wide, const float* rho, int nx, int ny, int nz, float sx,
                            float sy, float sz, float Q, float dr, int nsteps, float dx, float rcap, float bx,
                            float by, float bz, float* rfront, int nu, int nphi, void* stream) {
  clear_error();
  NV_REQUIRE(rho && rfront, "%s: null pointer", name);
  NV_REQUIRE(nx > 0 && ny > 0 && nz > 0 && nu > 0 && nphi > 0, "%s: bad sizes", name);
  NV_REQUIRE(dr > 0.0f && dx > 0.0f, "%s: dr and dx must be > 0", name);
  NV_REQUIRE((long long)nx * ny * nz < (1LL << 32), "%s: grids above 2^32 cells are not supported", name);
  const long long n = (long long)((nu + 7) / 8) * ((nphi + 7) / 8) * 64;
  const CellScale cs = cell_scale(dx);
  const bool pair = nz >= 2;
  auto kern = pair ? (wide ? cast_rays_kernel<true, true> : cast_rays_kernel<true, false>)
                   : (wide ? cast_rays_kernel<false, true> : cast_rays_kernel<false, false>);
  hipLaunchKernelGGL(kern, dim3(blocks_for(n, 256)), dim3(256), 0, as_stream(stream), rho, nx, ny, nz, sx, sy, sz, Q,
                     dr, nsteps, cs, dx, rcap, bx, by, bz, rfront, nu, nphi);
  return check_launch(name);
}

extern "C" int nv_cast_rays(const float* rho, int nx, int ny, int nz, float sx, float sy, float sz, float Q, float dr,
                            int nsteps, float dx, float rcap, float bx, float by, float bz, float* rfront, int nu,
                            int nphi, void* stream) {
  return cast_rays_launch("nv_cast_rays", false, rho, nx, ny, nz, sx, sy, sz, Q, dr, nsteps, dx, rcap, bx, by, bz,
                          rfront, nu, nphi, stream);
}

extern "C" int nv_cast_rays_wide(const float* rho, int nx, int ny, int nz, float sx, float sy, float sz, float Q,
                                 float dr, int nsteps, float dx, float rcap, float bx, float by, float bz,
                                 float* rfront, int nu, int nphi, void* stream) {
  return cast_rays_launch("nv_cast_rays_wide", true, rho, nx, ny, nz, sx, sy, sz, Q, dr, nsteps, dx, rcap, bx, by,
                          bz, rfront, nu, nphi, stream);
}

extern "C" int nv_apply_front(const float* rfront, int nu, int nphi, float sx, float sy, float sz, float dx,
                              float width, float* xi, float* dfront, int nx, int ny, int nz, void* stream) {
  clear_error();
  NV_REQUIRE(rfront && xi && dfront, "nv_apply_front: null pointer");
  NV_REQUIRE(nx > 0 && ny > 0 && nz > 0 && nu > 0 && nphi > 0, "nv_apply_front: bad sizes");
  const long long n = (long long)nx * ny * nz;
  hipLaunchKernelGGL(apply_front_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, as_stream(stream), rfront, nu, nphi,
                     sx, sy, sz, dx, width, xi, dfront, nx, ny, nz);
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

static int ionize_apply_launch(const char* name, bool vel, const float* rfronts, int nu, int nphi,
                               const float* star_xyz, int nstars, float dx, float width, float* xi, uint16_t* dfront_h,
                               float* dfront_f, int nx, int ny, int nz, const void* f0, const void* f1,
                               const void* f2, const void* f3, float cn2, float ci2, float* smax_out, void* stream) {
  clear_error();
  NV_REQUIRE(xi && (nstars == 0 || (rfronts && star_xyz)), "%s: null pointer", name);
  NV_REQUIRE((dfront_h != nullptr) != (dfront_f != nullptr), "%s: pass exactly one dfront buffer", name);
  NV_REQUIRE(nx > 0 && ny > 0 && nz > 0 && nu > 0 && nphi > 0 && nstars >= 0, "%s: bad sizes", name);
  NV_REQUIRE(!smax_out || (f0 && f1 && f2 && f3), "%s: smax needs the density and momentum fields", name);
  IonArgs A;
  A.rfronts = rfronts; A.star_xyz = star_xyz; A.xi = xi; A.df_h = dfront_h; A.df_f = dfront_f;
  A.f0 = f0; A.f1 = f1; A.f2 = f2; A.f3 = f3; A.smax = smax_out;
  A.nu = nu; A.nphi = nphi; A.nstars = nstars; A.nx = nx; A.ny = ny; A.nz = nz;
  A.dx = dx; A.width = width; A.cn2 = cn2; A.ci2 = ci2;
  const long long tiles_jk = (long long)((nz + IK - 1) / IK) * ((ny + IJ - 1) / IJ);
  A.ci = (int)std::min<long long>(64, std::max<long long>(4, (long long)nx * tiles_jk / 2048));
  dim3 grid((nz + IK - 1) / IK, (ny + IJ - 1) / IJ, (nx + A.ci - 1) / A.ci);
  auto kern = !smax_out ? ionize_apply_kernel<false, false>
                        : (vel ? ionize_apply_kernel<true, true> : ionize_apply_kernel<true, false>);
  hipLaunchKernelGGL(kern, grid, dim3(IK, IJ), 0, as_stream(stream), A);
  return check_launch(name);
}

extern "C" int nv_ionize_apply(const float* rfronts, int nu, int nphi, const float* star_xyz, int nstars, float dx,
                               float width, float* xi, uint16_t* dfront_h, float* dfront_f, int nx, int ny, int nz,
                               const float* rho, const float* mx, const float* my, const float* mz, float cn2,
                               float ci2, float* smax_out, void* stream) {
  return ionize_apply_launch("nv_ionize_apply", false, rfronts, nu, nphi, star_xyz, nstars, dx, width, xi, dfront_h,
                             dfront_f, nx, ny, nz, rho, mx, my, mz, cn2, ci2, smax_out, stream);
}

extern "C" int nv_ionize_apply_h(const float* rfronts, int nu, int nphi, const float* star_xyz, int nstars, float dx,
                                 float width, float* xi, uint16_t* dfront_h, float* dfront_f, int nx, int ny, int nz,
                                 const float* rho, const uint16_t* vx, const uint16_t* vy, const uint16_t* vz,
                                 float cn2, float ci2, float* smax_out, void* stream) {
  return ionize_apply_launch("nv_ionize_apply_h", true, rfronts, nu, nphi, star_xyz, nstars, dx, width, xi, dfront_h,
                             dfront_f, nx, ny, nz, rho, vx, vy, vz, cn2, ci2, smax_out, stream);
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
  const CellScale cs = cell_scale(dx);
  auto kern = nz >= 2 ? cast_fuv_kernel<true> : cast_fuv_kernel<false>;
  hipLaunchKernelGGL(kern, dim3(blocks_for(n, 256)), dim3(256), 0, as_stream(stream), rho, xi, nx, ny, nz, sx, sy, sz,
                     kappa, dr, nsteps, cs, binr, bx, by, bz, att, nr, nu, nphi);
  return check_launch("nv_cast_fuv");
}

extern "C" int nv_apply_fuv(const float* att, int nr, int nu, int nphi, float sx, float sy, float sz, float lum,
                            float dx, float binr, float* g0, int nx, int ny, int nz, void* stream) {
  clear_error();
  NV_REQUIRE(att && g0, "nv_apply_fuv: null pointer");
  NV_REQUIRE(nx > 0 && ny > 0 && nz > 0 && nr > 0 && nu > 0 && nphi > 0, "nv_apply_fuv: bad sizes");
  const long long n = (long long)nx * ny * nz;
  hipLaunchKernelGGL(apply_fuv_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, as_stream(stream), att, nr, nu, nphi,
                     sx, sy, sz, lum, dx, binr, g0, nx, ny, nz);
  return check_launch("nv_apply_fuv");
}

extern "C" int nv_fuv_apply(const float* atts, int nr, int nu, int nphi, const float* star_xyz, const float* lums,
                            int nstars, float dx, float binr, uint16_t* g0_h, float* g0_f, int nx, int ny, int nz,
                            void* stream) {
  clear_error();
  NV_REQUIRE(nstars == 0 || (atts && star_xyz && lums), "nv_fuv_apply: null pointer");
  NV_REQUIRE((g0_h != nullptr) != (g0_f != nullptr), "nv_fuv_apply: pass exactly one g0 buffer");
  NV_REQUIRE(nx > 0 && ny > 0 && nz > 0 && nr > 0 && nu > 0 && nphi > 0 && nstars >= 0, "nv_fuv_apply: bad sizes");
  const long long n = (long long)nx * ny * nz;
  hipLaunchKernelGGL(fuv_apply_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, as_stream(stream), atts, nr, nu, nphi,
                     star_xyz, lums, nstars, dx, binr, g0_h, g0_f, nx, ny, nz);
  return check_launch("nv_fuv_apply");
}
