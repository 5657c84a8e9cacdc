// nv_common.h - shared device helpers and the host-side error plumbing of the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string>
#include "../../include/nebula_hip.h"

namespace nv {

// ------------------------------------------------------------------------------------ host error plumbing
void set_error(const char* fmt, ...);
void clear_error();

#define NV_REQUIRE(cond, ...)        \
  do {                               \
    if (!(cond)) {                   \
      ::nv::set_error(__VA_ARGS__);  \
      return NV_ERR_ARG;             \
    }                                \
  } while (0)

// Checks the launch that was just issued; returns NV_OK or NV_ERR_LAUNCH with a message.
int check_launch(const char* what);

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// 1/x if x is a power of two whose reciprocal is a normal float (then a / x and a * (1/x) round the same real
// number, so they are equal bit for bit), else 0.
inline float pow2_reciprocal(float x) {
  int e = 0;
  const float m = frexpf(x, &e);
  if (!(x > 0.0f) || m != 0.5f || e < -124 || e > 126) return 0.0f;
  return ldexpf(1.0f, 1 - e);
}

// ------------------------------------------------------------------------------------ device helpers
// AMDGPU constant address space: for a table that no thread writes during the launch, loads at a wave-uniform
// index through such a pointer are scalar loads (no vector load and wait per lane)
#define CONST_AS __attribute__((address_space(4)))

// v clamped to [lo, hi] (lo <= hi at every call); one v_med3_i32
__device__ __forceinline__ int clampi(int v, int lo, int hi) { return min(max(v, lo), hi); }

// Workgroups are dealt to the 8 XCDs round robin (id % 8). A launch of 8 per_xcd workgroups that maps workgroup id to
// (id % 8) per_xcd + id / 8 gives each XCD a contiguous run of per_xcd work items, so items that share data (the
// halos of neighbouring tiles, the front-table rows of neighbouring directions) meet in one XCD's L2.
__device__ __forceinline__ int xcd_tile(int per_xcd) { return (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3); }

// C-order [nx][ny][nz] offset; 64-bit so 1024^3 grids (2^30 cells) index safely
__device__ __forceinline__ long long cell(int i, int j, int k, int ny, int nz) {
  return ((long long)i * ny + j) * (long long)nz + k;
}

// wave64 max of non-negative floats
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  return v;
}

// Block-wide max of non-negative floats into dst[0] (float bits compared as unsigned ints, which orders
// non-negative floats correctly). `red` is LDS scratch of at least blockDim/64 floats.
__device__ __forceinline__ void block_max_atomic(float v, float* red, float* dst) {
  const int nthreads = blockDim.x * blockDim.y * blockDim.z;
  const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
  v = wave_max(v);
  __syncthreads();
  if ((tid & 63) == 0) red[tid >> 6] = v;
  __syncthreads();
  if (tid == 0) {
    float m = red[0];
    for (int w = 1; w < (nthreads + 63) / 64; ++w) m = fmaxf(m, red[w]);
    atomicMax(reinterpret_cast<unsigned int*>(dst), __float_as_uint(m));
  }
}

__device__ __forceinline__ uint16_t f32_to_f16_bits(float x) {
  return __builtin_bit_cast(uint16_t, (_Float16)x);
}

}  // namespace nv
