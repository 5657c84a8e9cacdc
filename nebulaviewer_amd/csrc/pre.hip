// pre.hip - the pre-turbulence phase around the step (sim.py:1030-1060): turbulent stirring from two low-
// resolution solenoidal fields (nebula3d.drive, nebula3d.py:165-179), the torus added at the end of the phase
// (sim.add_torus, sim.py:352-363), the mass-weighted Mach sums (sim.mach_sums, sim.py:501-508, here in FP64),
// and the 2x subsampling the star-formation trigger reads (sim.every2 / every2_half, sim.py:517-527).
#include <stdint.h>
#include "nv_common.h"
#include "nv_divide.h"
#include "nv_sample.h"

namespace nv {

struct DriveArgs {
  float *rho, *mx, *my, *mz;       // rho read, m updated in place
  const float *fx, *fy, *fz, *gx, *gy, *gz;
  int nx, ny, nz, sx, sy, sz;      // grid and stirring-grid shapes
  float blend, amp, dt, dx;
};

// a = amp * lerp(F, G, blend) at each cell centre, F and G sampled trilinearly on the stirring grid; m += rho amp
// dt * lerp (nebula3d.py:170-179). The stirring grid is small (24 cells per unit plus guards): paired z loads and
// 32-bit offsets.
template <class SC>
__global__ __launch_bounds__(256) void drive_kernel(DriveArgs A, SC cs) {
  const long long n = (long long)A.nx * A.ny * A.nz;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) {
    const int k = (int)(q % A.nz);
    const long long ij = q / A.nz;
    const int j = (int)(ij % A.ny), i = (int)(ij / A.ny);
    const float px = ((float)i + 0.5f) * A.dx, py = ((float)j + 0.5f) * A.dx, pz = ((float)k + 0.5f) * A.dx;
    float f[3];
    const float* Fa[3] = {A.fx, A.fy, A.fz};
    const float* Fb[3] = {A.gx, A.gy, A.gz};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float fa = sample_r<true, uint32_t>(Fa[c], A.sx, A.sy, A.sz, px, py, pz, cs);
      const float fb = sample_r<true, uint32_t>(Fb[c], A.sx, A.sy, A.sz, px, py, pz, cs);
      f[c] = (1.0f - A.blend) * fa + A.blend * fb;
    }
    const float r = A.rho[q] * A.amp * A.dt;
    A.mx[q] = A.mx[q] + r * f[0];
    A.my[q] = A.my[q] + r * f[1];
    A.mz[q] = A.mz[q] + r * f[2];
  }
}

// sim.add_torus (sim.py:353-363)
__global__ __launch_bounds__(256) void add_torus_kernel(float* rho, float* mx, float* my, float* mz,
                                                        const float* __restrict__ torus, long long n) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) {
    const float r = rho[q], tq = torus[q];
    const float f = 1.0f - tq / (tq + r);
    rho[q] = r + tq;
    mx[q] = mx[q] * f;
    my[q] = my[q] * f;
    mz[q] = mz[q] * f;
  }
}

// out[0] += sum |m|^2 / rho, out[1] += sum rho (per cell in float as sim.py:505-508, summed in FP64)
__global__ __launch_bounds__(256) void mach_sums_kernel(const float* __restrict__ rho, const float* __restrict__ mx,
                                                        const float* __restrict__ my, const float* __restrict__ mz,
                                                        long long n, double* out) {
  __shared__ double red[2][4];
  double a = 0.0, b = 0.0;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) {
    const float r = rho[q];
    const float m2 = (mx[q] * mx[q] + my[q] * my[q]) + mz[q] * mz[q];
    a += (double)(m2 / r);
    b += (double)r;
  }
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_down(a, o, 64);
    b += __shfl_down(b, o, 64);
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    red[0][w] = a;
    red[1][w] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double sa = 0.0, sb = 0.0;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) {
      sa += red[0][k];
      sb += red[1][k];
    }
    atomicAdd(&out[0], sa);
    atomicAdd(&out[1], sb);
  }
}

// dst = src[::2, ::2, ::2] (sim.every2 / every2_half, sim.py:517-527); HALF: binary16 source
template <bool HALF>
__global__ __launch_bounds__(256) void every2_kernel(const void* __restrict__ src, float* __restrict__ dst, int nx,
                                                     int ny, int nz) {
  const int hx = (nx + 1) / 2, hy = (ny + 1) / 2, hz = (nz + 1) / 2;
  const long long n = (long long)hx * hy * hz;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) {
    const int k = (int)(q % hz);
    const long long ij = q / hz;
    const int j = (int)(ij % hy), i = (int)(ij / hy);
    const long long s = cell(2 * i, 2 * j, 2 * k, ny, nz);
    if constexpr (HALF) dst[q] = (float)__builtin_bit_cast(_Float16, static_cast<const uint16_t*>(src)[s]);
    else dst[q] = static_cast<const float*>(src)[s];
  }
}

inline int grid_of(long long n) { return (int)std::min<long long>((n + 255) / 256, 65536); }

}  // namespace nv

using namespace nv;

extern "C" int nv_drive(float* rho, float* mx, float* my, float* mz, int nx, int ny, int nz, const float* fx,
                        const float* fy, const float* fz, const float* gx, const float* gy, const float* gz, int sx,
                        int sy, int sz, float blend, float amp, float dt, float dx, float dxs, void* stream) {
  clear_error();
  NV_REQUIRE(rho && mx && my && mz && fx && fy && fz && gx && gy && gz, "nv_drive: null pointer");
  NV_REQUIRE(nx > 0 && ny > 0 && nz > 0 && sx > 0 && sy > 0 && sz >= 2,
             "nv_drive: bad sizes (the stirring grid needs sz >= 2)");
  NV_REQUIRE((long long)sx * sy * sz <= (1LL << 30) && (long long)sx * sy < (1LL << 24),
             "nv_drive: stirring grid above 2^30 cells or 2^24 rows");
  NV_REQUIRE(dx > 0.0f && dxs >= 1.17549435e-38f && dxs <= 3.40282347e+38f, "nv_drive: dx, dxs must be > 0");
  DriveArgs A{rho, mx, my, mz, fx, fy, fz, gx, gy, gz, nx, ny, nz, sx, sy, sz, blend, amp, dt, dx};
  const long long n = (long long)nx * ny * nz;
  const float inv = pow2_reciprocal(dxs);
  if (inv != 0.0f) {
    ScalePow2 sc;
    sc.inv = inv;
    hipLaunchKernelGGL(drive_kernel<ScalePow2>, dim3(grid_of(n)), dim3(256), 0, as_stream(stream), A, sc);
  } else {
    ScaleDiv sc;
    sc.rdx = Recip{1.0 / (double)dxs, dxs};
    hipLaunchKernelGGL(drive_kernel<ScaleDiv>, dim3(grid_of(n)), dim3(256), 0, as_stream(stream), A, sc);
  }
  return check_launch("nv_drive");
}

extern "C" int nv_add_torus(float* rho, float* mx, float* my, float* mz, const float* torus, long long n,
                            void* stream) {
  clear_error();
  NV_REQUIRE(rho && mx && my && mz && torus && n >= 0, "nv_add_torus: bad arguments");
  if (n == 0) return NV_OK;
  hipLaunchKernelGGL(add_torus_kernel, dim3(grid_of(n)), dim3(256), 0, as_stream(stream), rho, mx, my, mz, torus, n);
  return check_launch("nv_add_torus");
}

extern "C" int nv_mach_sums(const float* rho, const float* mx, const float* my, const float* mz, long long n,
                            double* out, void* stream) {
  clear_error();
  NV_REQUIRE(rho && mx && my && mz && out && n >= 0, "nv_mach_sums: bad arguments");
  if (n == 0) return NV_OK;
  const int blocks = (int)std::min<long long>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(mach_sums_kernel, dim3(blocks), dim3(256), 0, as_stream(stream), rho, mx, my, mz, n, out);
  return check_launch("nv_mach_sums");
}

extern "C" int nv_every2(const void* src, int src_half, float* dst, int nx, int ny, int nz, void* stream) {
  clear_error();
  NV_REQUIRE(src && dst && nx > 0 && ny > 0 && nz > 0, "nv_every2: bad arguments");
  const long long n = (long long)((nx + 1) / 2) * ((ny + 1) / 2) * ((nz + 1) / 2);
  if (src_half)
    hipLaunchKernelGGL(every2_kernel<true>, dim3(grid_of(n)), dim3(256), 0, as_stream(stream), src, dst, nx, ny, nz);
  else
    hipLaunchKernelGGL(every2_kernel<false>, dim3(grid_of(n)), dim3(256), 0, as_stream(stream), src, dst, nx, ny, nz);
  return check_launch("nv_every2");
}
