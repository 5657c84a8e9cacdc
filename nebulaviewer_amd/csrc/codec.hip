// codec.hip - the kernels consumers run on the step's fields, next to the step itself:
//   timeline snapshot codec   k_pack / k_pack_lm / k_unpack / k_unpack_lm (timeline.py:19-89), 13 B per cell
//   shock proxy               shock_field (nebula3d.py:182-191)
//   bake frame stored forms   the numpy conversions of bake.py:94-99 (fp16 / uint8 arrays written per frame)
//   render volume packing     pack_volume_lm and smooth_axis (render.py:176-204, 281-298)
// All are bandwidth- or issue-bound elementwise passes and short stencils over the SoA fields: the codec moves
// four cells per thread with 16-byte float loads and 8-byte binary16 stores; divisions by a constant go through
// the shared double reciprocal (nv_divide.h), which gives the IEEE quotient.
#include <math.h>
#include <stdint.h>
#include "nv_common.h"
#include "nv_divide.h"

namespace nv {

__device__ __forceinline__ float h2f_(uint16_t h) { return (float)__builtin_bit_cast(_Float16, h); }

// numpy-style saturations: a NaN input stays NaN (np.minimum / np.clip propagate it)
__device__ __forceinline__ float min_np(float x, float m) { return x > m ? m : x; }
__device__ __forceinline__ float clip_np(float x, float lo, float hi) { return x < lo ? lo : (x > hi ? hi : x); }
// Warp's wp.min / wp.clamp on finite values
__device__ __forceinline__ float wmin(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float wclamp(float x, float lo, float hi) { return wmin(x > lo ? x : lo, hi); }

template <bool HALF>
__device__ __forceinline__ float ld_field(const void* p, long long q) {
  if constexpr (HALF) return h2f_(static_cast<const uint16_t*>(p)[q]);
  else return static_cast<const float*>(p)[q];
}

template <bool HALF>
__device__ __forceinline__ void st_field(void* p, long long q, uint16_t h) {
  if constexpr (HALF) static_cast<uint16_t*>(p)[q] = h;
  else static_cast<float*>(p)[q] = h2f_(h);
}

// ------------------------------------------------------------------------------------ timeline codec
struct PackArgs {
  const float *rho, *mx, *my, *mz, *xi;
  const void *df, *g0;
  uint16_t *o_r, *o_mx, *o_my, *o_mz, *o_df, *o_g0;
  uint8_t* o_xi;
  long long n;
};

// one cell of k_pack / k_pack_lm (timeline.py:23-34, 61-72)
template <bool HALF>
__device__ __forceinline__ void pack_cell(const PackArgs& A, long long q) {
  const float m = 65000.0f;
  A.o_r[q] = f32_to_f16_bits(wmin(A.rho[q], m));
  A.o_mx[q] = f32_to_f16_bits(wclamp(A.mx[q], -m, m));
  A.o_my[q] = f32_to_f16_bits(wclamp(A.my[q], -m, m));
  A.o_mz[q] = f32_to_f16_bits(wclamp(A.mz[q], -m, m));
  A.o_xi[q] = (uint8_t)(int)(wclamp(A.xi[q], 0.0f, 1.0f) * 255.0f + 0.5f);
  A.o_df[q] = f32_to_f16_bits(wclamp(ld_field<HALF>(A.df, q), -m, m));
  A.o_g0[q] = f32_to_f16_bits(wmin(ld_field<HALF>(A.g0, q), m));
}

template <bool HALF>
__global__ __launch_bounds__(256) void timeline_pack_kernel(PackArgs A) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < A.n; q += stride) pack_cell<HALF>(A, q);
}

struct UnpackArgs {
  const uint16_t *s_r, *s_mx, *s_my, *s_mz, *s_df, *s_g0;
  const uint8_t* s_xi;
  float floor_;
  float *rho, *mx, *my, *mz, *xi;
  void *df, *g0;
  long long n;
};

constexpr Recip R255 = {1.0 / 255.0, 255.0f};   // div_r(x, R255) is x / 255.0f

// k_unpack / k_unpack_lm (timeline.py:46-52, 84-89)
template <bool HALF>
__global__ __launch_bounds__(256) void timeline_unpack_kernel(UnpackArgs A) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < A.n; q += stride) {
    const float r = h2f_(A.s_r[q]);
    A.rho[q] = r > A.floor_ ? r : A.floor_;
    A.mx[q] = h2f_(A.s_mx[q]);
    A.my[q] = h2f_(A.s_my[q]);
    A.mz[q] = h2f_(A.s_mz[q]);
    A.xi[q] = div_r((float)(int)A.s_xi[q], R255);
    st_field<HALF>(A.df, q, A.s_df[q]);
    st_field<HALF>(A.g0, q, A.s_g0[q]);
  }
}

// ------------------------------------------------------------------------------------ shock proxy
// max(0, -div v) rho with central differences of v = m / rho over clamped neighbours (nebula3d.py:187-191);
// r2dx = 1 / (2 dx) as the shared reciprocal of the float 2 dx
__device__ __forceinline__ float shock_cell(const float* __restrict__ rho, const float* __restrict__ mx,
                                            const float* __restrict__ my, const float* __restrict__ mz, int i, int j,
                                            int k, int nx, int ny, int nz, const Recip& r2dx) {
  const int ip = min(i + 1, nx - 1), im = max(i - 1, 0), jp = min(j + 1, ny - 1), jm = max(j - 1, 0);
  const int kp = min(k + 1, nz - 1), km = max(k - 1, 0);
  const long long a = cell(ip, j, k, ny, nz), b = cell(im, j, k, ny, nz);
  const long long c = cell(i, jp, k, ny, nz), d = cell(i, jm, k, ny, nz);
  const long long e = cell(i, j, kp, ny, nz), f = cell(i, j, km, ny, nz);
  const float du = mx[a] / rho[a] - mx[b] / rho[b];
  const float dv = my[c] / rho[c] - my[d] / rho[d];
  const float dw = mz[e] / rho[e] - mz[f] / rho[f];
  const float div = div_r((du + dv) + dw, r2dx);
  const float s = -div;
  return (s > 0.0f ? s : 0.0f) * rho[cell(i, j, k, ny, nz)];
}

__global__ __launch_bounds__(256) void shock_kernel(const float* __restrict__ rho, const float* __restrict__ mx,
                                                    const float* __restrict__ my, const float* __restrict__ mz,
                                                    int nx, int ny, int nz, Recip r2dx, float* __restrict__ out) {
  const long long n = (long long)nx * ny * nz;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) {
    const int k = (int)(q % nz);
    const long long ij = q / nz;
    const int j = (int)(ij % ny), i = (int)(ij / ny);
    out[q] = shock_cell(rho, mx, my, mz, i, j, k, nx, ny, nz, r2dx);
  }
}

// ------------------------------------------------------------------------------------ bake frame forms
struct FrameArgs {
  const float *rho, *xi, *sh;
  const void *df, *g0;
  uint16_t *o_rho, *o_df, *o_g0, *o_sh;
  uint8_t* o_xi;
  long long n;
};

// bake.py:94-99: g0 -> min(g0, 6e4) fp16; rho -> fp16; xi -> round(clip(xi, 0, 1) * 255) uint8 (numpy rounds
// half to even); dfront -> clip(df, -4, 4) fp16; shock -> min(sh, 6e4) fp16
template <bool HALF>
__global__ __launch_bounds__(256) void frame_pack_kernel(FrameArgs A) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < A.n; q += stride) {
    A.o_g0[q] = f32_to_f16_bits(min_np(ld_field<HALF>(A.g0, q), 6e4f));
    A.o_rho[q] = f32_to_f16_bits(A.rho[q]);
    A.o_xi[q] = (uint8_t)__builtin_rintf(clip_np(A.xi[q], 0.0f, 1.0f) * 255.0f);
    A.o_df[q] = f32_to_f16_bits(clip_np(ld_field<HALF>(A.df, q), -4.0f, 4.0f));
    if (A.sh) A.o_sh[q] = f32_to_f16_bits(min_np(A.sh[q], 6e4f));
  }
}

// ------------------------------------------------------------------------------------ render packing
struct VolArgs {
  const float *rho, *mx, *my, *mz, *xi;
  const void *df, *g0;
  NvPackParams pp;
  Recip r_gs, r_edge, r2dx;        // gs, edge (r = 0 if edge <= 0) and 2 dx with their shared reciprocals
  int nx, ny, nz, shock_on;
  uint16_t* out;                   // binary16 x 4 per cell
};

// sstep(e0, e1, x) of render.py:121-124 with e1 - e0 = d and its shared reciprocal rd
__device__ __forceinline__ float sstep_r(float e0, const Recip& rd, float x) {
  const float t = wclamp(div_r(x - e0, rd), 0.0f, 1.0f);
  return t * t * (3.0f - 2.0f * t);
}

// fade_axis(u, e) of render.py:127-130 (rde = 1/e, e > 0)
__device__ __forceinline__ float fade_r(float u, const Recip& rde) { return sstep_r(0.0f, rde, u) * sstep_r(0.0f, rde, 1.0f - u); }

constexpr float OIII_E0 = -0.03f, OIII_D = -0.12f - (-0.03f);   // sstep(-0.03, -0.12, df)
constexpr float SII_E0 = -0.05f, SII_D = -0.005f - (-0.05f);    // sstep(-0.05, -0.005, df)
constexpr Recip R_OIII = {1.0 / (double)OIII_D, OIII_D}, R_SII = {1.0 / (double)SII_D, SII_D};

template <bool HALF>
__global__ __launch_bounds__(256) void pack_volume_kernel(VolArgs A) {
  const long long n = (long long)A.nx * A.ny * A.nz;
  const long long stride = (long long)gridDim.x * blockDim.x;
  const NvPackParams& pp = A.pp;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) {
    const int k = (int)(q % A.nz);
    const long long ij = q / A.nz;
    const int j = (int)(ij % A.ny), i = (int)(ij / A.ny);
    const float r = A.rho[q];
    const float x = wclamp(A.xi[q], 0.0f, 1.0f);
    const float d = ld_field<HALF>(A.df, q);
    const float ni = r * x;
    const float ha = ni * ni;
    const float oiii = ha * sstep_r(OIII_E0, R_OIII, d);
    const float sii = ha * sstep_r(SII_E0, R_SII, d);
    float dust = r * (1.0f - x) + 0.08f * ni;
    const float g = ld_field<HALF>(A.g0, q);
    const float geff = g / (1.0f + div_r(g, A.r_gs)) + pp.ambient;
    const float sh = A.shock_on ? shock_cell(A.rho, A.mx, A.my, A.mz, i, j, k, A.nx, A.ny, A.nz, A.r2dx) : 0.0f;
    const float p0 = (((float)i + 0.5f) * pp.dx) * pp.inv_L[0];
    const float p1 = (((float)j + 0.5f) * pp.dx) * pp.inv_L[1];
    const float p2 = (((float)k + 0.5f) * pp.dx) * pp.inv_L[2];
    const float f = A.r_edge.r != 0.0 ? (fade_r(p0, A.r_edge) * fade_r(p1, A.r_edge)) * fade_r(p2, A.r_edge) : 1.0f;
    const float dg = dust * geff;
    float em[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      em[c] = (((ha * pp.c_ha[c] + oiii * pp.c_o3[c]) + sii * pp.c_s2[c]) + dg * pp.c_dg[c]) + sh * pp.c_sh[c];
      em[c] = em[c] * f;
    }
    dust = dust * f;
    const float m = 60000.0f;
    typedef uint16_t h4 __attribute__((ext_vector_type(4)));
    h4 o;
    o.x = f32_to_f16_bits(wmin(em[0], m));
    o.y = f32_to_f16_bits(wmin(em[1], m));
    o.z = f32_to_f16_bits(wmin(em[2], m));
    o.w = f32_to_f16_bits(wmin(dust, m));
    reinterpret_cast<h4*>(A.out)[q] = o;
  }
}

constexpr Recip R6C = {1.0 / 6.0, 6.0f};

// smooth_axis (render.py:281-298): [1 4 1]/6 along one axis of a binary16 x 4 grid, clamped at the ends
__global__ __launch_bounds__(256) void smooth_axis_kernel(const uint16_t* __restrict__ src, uint16_t* __restrict__ dst,
                                                          int axis, int nx, int ny, int nz) {
  typedef uint16_t h4 __attribute__((ext_vector_type(4)));
  const long long n = (long long)nx * ny * nz;
  const long long stride = (long long)gridDim.x * blockDim.x;
  const h4* s = reinterpret_cast<const h4*>(src);
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) {
    const int k = (int)(q % nz);
    const long long ij = q / nz;
    const int j = (int)(ij % ny), i = (int)(ij / ny);
    long long qa, qc;
    if (axis == 0) {
      qa = cell(max(i - 1, 0), j, k, ny, nz);
      qc = cell(min(i + 1, nx - 1), j, k, ny, nz);
    } else if (axis == 1) {
      qa = cell(i, max(j - 1, 0), k, ny, nz);
      qc = cell(i, min(j + 1, ny - 1), k, ny, nz);
    } else {
      qa = cell(i, j, max(k - 1, 0), ny, nz);
      qc = cell(i, j, min(k + 1, nz - 1), ny, nz);
    }
    const h4 a = s[qa], b = s[q], c = s[qc];
    h4 o;
    o.x = f32_to_f16_bits(div_r((h2f_(a.x) + 4.0f * h2f_(b.x)) + h2f_(c.x), R6C));
    o.y = f32_to_f16_bits(div_r((h2f_(a.y) + 4.0f * h2f_(b.y)) + h2f_(c.y), R6C));
    o.z = f32_to_f16_bits(div_r((h2f_(a.z) + 4.0f * h2f_(b.z)) + h2f_(c.z), R6C));
    o.w = f32_to_f16_bits(div_r((h2f_(a.w) + 4.0f * h2f_(b.w)) + h2f_(c.w), R6C));
    reinterpret_cast<h4*>(dst)[q] = o;
  }
}

inline int grid_for(long long n) { return (int)std::min<long long>((n + 255) / 256, 65536); }

inline bool normal_pos(float x) { return x >= 1.17549435e-38f && x <= 3.40282347e+38f; }

}  // namespace nv

using namespace nv;

extern "C" int nv_timeline_pack(const float* rho, const float* mx, const float* my, const float* mz, const float* xi,
                                const void* dfront, const void* g0, int fields_half, uint16_t* o_r, uint16_t* o_mx,
                                uint16_t* o_my, uint16_t* o_mz, uint8_t* o_xi, uint16_t* o_df, uint16_t* o_g0,
                                long long n, void* stream) {
  clear_error();
  NV_REQUIRE(rho && mx && my && mz && xi && dfront && g0 && o_r && o_mx && o_my && o_mz && o_xi && o_df && o_g0,
             "nv_timeline_pack: null pointer");
  NV_REQUIRE(n >= 0, "nv_timeline_pack: n < 0");
  if (n == 0) return NV_OK;
  PackArgs A{rho, mx, my, mz, xi, dfront, g0, o_r, o_mx, o_my, o_mz, o_df, o_g0, o_xi, n};
  if (fields_half)
    hipLaunchKernelGGL(timeline_pack_kernel<true>, dim3(grid_for(n)), dim3(256), 0, as_stream(stream), A);
  else
    hipLaunchKernelGGL(timeline_pack_kernel<false>, dim3(grid_for(n)), dim3(256), 0, as_stream(stream), A);
  return check_launch("nv_timeline_pack");
}

extern "C" int nv_timeline_unpack(const uint16_t* s_r, const uint16_t* s_mx, const uint16_t* s_my,
                                  const uint16_t* s_mz, const uint8_t* s_xi, const uint16_t* s_df,
                                  const uint16_t* s_g0, float floor_, float* rho, float* mx, float* my, float* mz,
                                  float* xi, void* dfront, void* g0, int fields_half, long long n, void* stream) {
  clear_error();
  NV_REQUIRE(s_r && s_mx && s_my && s_mz && s_xi && s_df && s_g0 && rho && mx && my && mz && xi && dfront && g0,
             "nv_timeline_unpack: null pointer");
  NV_REQUIRE(n >= 0, "nv_timeline_unpack: n < 0");
  if (n == 0) return NV_OK;
  UnpackArgs A{s_r, s_mx, s_my, s_mz, s_df, s_g0, s_xi, floor_, rho, mx, my, mz, xi, dfront, g0, n};
  if (fields_half)
    hipLaunchKernelGGL(timeline_unpack_kernel<true>, dim3(grid_for(n)), dim3(256), 0, as_stream(stream), A);
  else
    hipLaunchKernelGGL(timeline_unpack_kernel<false>, dim3(grid_for(n)), dim3(256), 0, as_stream(stream), A);
  return check_launch("nv_timeline_unpack");
}

extern "C" int nv_shock_field(const float* rho, const float* mx, const float* my, const float* mz, int nx, int ny,
                              int nz, float dx, float* out, void* stream) {
  clear_error();
  NV_REQUIRE(rho && mx && my && mz && out, "nv_shock_field: null pointer");
  NV_REQUIRE(nx > 0 && ny > 0 && nz > 0, "nv_shock_field: bad grid");
  NV_REQUIRE(normal_pos(2.0f * dx), "nv_shock_field: dx must be a positive normal float");
  const long long n = (long long)nx * ny * nz;
  hipLaunchKernelGGL(shock_kernel, dim3(grid_for(n)), dim3(256), 0, as_stream(stream), rho, mx, my, mz, nx, ny, nz,
                     Recip{1.0 / (double)(2.0f * dx), 2.0f * dx}, out);
  return check_launch("nv_shock_field");
}

extern "C" int nv_frame_pack(const float* rho, const float* xi, const void* dfront, const void* g0, const float* shock,
                             int fields_half, uint16_t* o_rho, uint8_t* o_xi, uint16_t* o_df, uint16_t* o_g0,
                             uint16_t* o_sh, long long n, void* stream) {
  clear_error();
  NV_REQUIRE(rho && xi && dfront && g0 && o_rho && o_xi && o_df && o_g0, "nv_frame_pack: null pointer");
  NV_REQUIRE((shock == nullptr) == (o_sh == nullptr), "nv_frame_pack: shock and o_sh go together");
  NV_REQUIRE(n >= 0, "nv_frame_pack: n < 0");
  if (n == 0) return NV_OK;
  FrameArgs A{rho, xi, shock, dfront, g0, o_rho, o_df, o_g0, o_sh, o_xi, n};
  if (fields_half)
    hipLaunchKernelGGL(frame_pack_kernel<true>, dim3(grid_for(n)), dim3(256), 0, as_stream(stream), A);
  else
    hipLaunchKernelGGL(frame_pack_kernel<false>, dim3(grid_for(n)), dim3(256), 0, as_stream(stream), A);
  return check_launch("nv_frame_pack");
}

extern "C" int nv_pack_volume(const float* rho, const float* mx, const float* my, const float* mz, const float* xi,
                              const void* dfront, const void* g0, int fields_half, const NvPackParams* pp,
                              int shock_on, uint16_t* out, int nx, int ny, int nz, void* stream) {
  clear_error();
  NV_REQUIRE(rho && xi && dfront && g0 && pp && out, "nv_pack_volume: null pointer");
  NV_REQUIRE(!shock_on || (mx && my && mz), "nv_pack_volume: the shock term needs the momentum fields");
  NV_REQUIRE(nx > 0 && ny > 0 && nz > 0, "nv_pack_volume: bad grid");
  NV_REQUIRE(normal_pos(pp->gs), "nv_pack_volume: gs must be a positive normal float");
  NV_REQUIRE(pp->edge <= 0.0f || normal_pos(pp->edge), "nv_pack_volume: edge must be <= 0 or a normal float");
  NV_REQUIRE(normal_pos(2.0f * pp->dx), "nv_pack_volume: dx must be a positive normal float");
  VolArgs A;
  A.rho = rho; A.mx = mx; A.my = my; A.mz = mz; A.xi = xi; A.df = dfront; A.g0 = g0;
  A.pp = *pp;
  A.r_gs = Recip{1.0 / (double)pp->gs, pp->gs};
  A.r_edge = pp->edge > 0.0f ? Recip{1.0 / (double)pp->edge, pp->edge} : Recip{0.0, 0.0f};
  A.r2dx = Recip{1.0 / (double)(2.0f * pp->dx), 2.0f * pp->dx};
  A.nx = nx; A.ny = ny; A.nz = nz; A.shock_on = shock_on ? 1 : 0;
  A.out = out;
  const long long n = (long long)nx * ny * nz;
  if (fields_half)
    hipLaunchKernelGGL(pack_volume_kernel<true>, dim3(grid_for(n)), dim3(256), 0, as_stream(stream), A);
  else
    hipLaunchKernelGGL(pack_volume_kernel<false>, dim3(grid_for(n)), dim3(256), 0, as_stream(stream), A);
  return check_launch("nv_pack_volume");
}

extern "C" int nv_smooth_axis(const uint16_t* src, uint16_t* dst, int axis, int nx, int ny, int nz, void* stream) {
  clear_error();
  NV_REQUIRE(src && dst && src != dst, "nv_smooth_axis: null or aliased pointers");
  NV_REQUIRE(axis >= 0 && axis <= 2, "nv_smooth_axis: axis must be 0, 1 or 2");
  NV_REQUIRE(nx > 0 && ny > 0 && nz > 0, "nv_smooth_axis: bad grid");
  const long long n = (long long)nx * ny * nz;
  hipLaunchKernelGGL(smooth_axis_kernel, dim3(grid_for(n)), dim3(256), 0, as_stream(stream), src, dst, axis, nx, ny,
                     nz);
  return check_launch("nv_smooth_axis");
}
