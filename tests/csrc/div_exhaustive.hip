// Exhaustive checks of the float-only exact division and square root of nebulaviewer_amd/csrc/nv_divide.h on the
// GPU it runs on
// (test helper, built into tests/csrc/libnv_divcheck.so):
//   nv_check_recip_f:  recip_f(b).y == RN(1/b) for every normal b below 2^126 (exponent fields 1..252);
//   nv_check_div_f:    div_fn == the IEEE quotient for all 2^46 pairs of significands a, b in [1, 2) (every
//                      astep-th a), the reference being the double-reciprocal quotient, which is correctly
//                      rounded for normal quotients;
//   nv_check_sqrt_rn:  sqrt_rn == the IEEE square root for every float from 2^-96 through +inf;
//   nv_div_fn_apply, nv_sqrt_rn_apply: the same functions on caller-chosen operands (the guards' edge cases).
// Each returns the number of mismatches (or -1 on a HIP error) and record up to 16 offending operand pairs.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../nebulaviewer_amd/csrc/nv_divide.h"

namespace {

__global__ __launch_bounds__(256) void recip_kernel(uint32_t ebits, unsigned long long* bad, uint32_t* rec) {
  const uint32_t m = blockIdx.x * 256u + threadIdx.x;   // grid of 2^15 blocks: m < 2^23
  const float b = __uint_as_float((ebits << 23) | m);
  const float y = nv::recip_f(b).y;
  if (y != (float)(1.0 / (double)b)) {
    atomicAdd(bad, 1ull);
    const uint32_t k = atomicAdd(&rec[0], 1u);
    if (k < 16) {
      rec[1 + 2 * k] = __float_as_uint(b);
      rec[2 + 2 * k] = __float_as_uint(y);
    }
  }
}

__global__ __launch_bounds__(256) void pairs_kernel(uint32_t bstart, uint32_t astep, unsigned long long* bad,
                                                    uint32_t* rec) {
  const uint32_t bm = bstart + blockIdx.x * 256u + threadIdx.x;   // the host keeps bm < 2^23
  const float b = __uint_as_float(0x3f800000u | bm);
  const nv::RecipF R = nv::recip_f(b);
  const double rd = nv::recip_d(b);
  unsigned long long nb = 0;
  for (uint32_t am = 0; am < (1u << 23); am += astep) {
    const float x[1] = {__uint_as_float(0x3f800000u | am)};
    float q[1];
    nv::div_fn<1>(x, R, q);
    if (q[0] != (float)((double)x[0] * rd)) {
      ++nb;
      const uint32_t k = atomicAdd(&rec[0], 1u);
      if (k < 16) {
        rec[1 + 2 * k] = __float_as_uint(x[0]);
        rec[2 + 2 * k] = __float_as_uint(b);
      }
    }
  }
  if (nb) atomicAdd(bad, nb);
}

// sqrt_rn == IEEE sqrtf for every float with exponent field e (e >= 31: from 2^-96 up, through +inf)
__global__ __launch_bounds__(256) void sqrt_kernel(uint32_t ebits, unsigned long long* bad, uint32_t* rec) {
  const uint32_t m = blockIdx.x * 256u + threadIdx.x;   // grid of 2^15 blocks: m < 2^23
  const uint32_t bits = (ebits << 23) | m;
  if (ebits == 255u && m != 0u) return;                  // NaN payloads
  const float x = __uint_as_float(bits);
  const float s = nv::sqrt_rn(x);
  const float ref = (float)sqrt((double)x);              // double rounding is innocuous for sqrt (53 >= 2*24 + 2)
  if (__float_as_uint(s) != __float_as_uint(ref)) {
    atomicAdd(bad, 1ull);
    const uint32_t k = atomicAdd(&rec[0], 1u);
    if (k < 16) {
      rec[1 + 2 * k] = bits;
      rec[2 + 2 * k] = __float_as_uint(s);
    }
  }
}

// q[i] = div_fn over groups of G numerators a[g G .. g G + G-1] sharing the divisor b[g] (device pointers)
template <int G>
__global__ __launch_bounds__(256) void apply_kernel(const float* __restrict__ a, const float* __restrict__ b,
                                                    float* __restrict__ q, int groups) {
  const int g = blockIdx.x * 256 + threadIdx.x;
  if (g >= groups) return;
  float x[G], y[G];
#pragma unroll
  for (int i = 0; i < G; ++i) x[i] = a[g * G + i];
  nv::div_fn<G>(x, nv::recip_f(b[g]), y);
#pragma unroll
  for (int i = 0; i < G; ++i) q[g * G + i] = y[i];
}

// q[i] = sqrt_rn(x[i]) (device pointers)
__global__ __launch_bounds__(256) void sqrt_apply_kernel(const float* __restrict__ x, float* __restrict__ s, int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) s[i] = nv::sqrt_rn(x[i]);
}

struct Buf {
  unsigned long long* bad = nullptr;
  uint32_t* rec = nullptr;
  bool ok() const { return bad && rec; }
  Buf() {
    if (hipMalloc(&bad, sizeof(unsigned long long)) != hipSuccess) bad = nullptr;
    if (hipMalloc(&rec, 33 * sizeof(uint32_t)) != hipSuccess) rec = nullptr;
    if (ok() && (hipMemset(bad, 0, sizeof(unsigned long long)) != hipSuccess ||
                 hipMemset(rec, 0, 33 * sizeof(uint32_t)) != hipSuccess)) {
      hipFree(bad);
      hipFree(rec);
      bad = nullptr;
      rec = nullptr;
    }
  }
  ~Buf() {
    if (bad) hipFree(bad);
    if (rec) hipFree(rec);
  }
  long long finish(uint32_t* out_rec) {
    unsigned long long n = 0;
    if (hipDeviceSynchronize() != hipSuccess) return -1;
    if (hipMemcpy(&n, bad, sizeof(n), hipMemcpyDeviceToHost) != hipSuccess) return -1;
    if (out_rec && hipMemcpy(out_rec, rec, 33 * sizeof(uint32_t), hipMemcpyDeviceToHost) != hipSuccess) return -1;
    return (long long)n;
  }
};

}  // namespace

// out_rec: 33 uint32 (count, then pairs (b, y)), may be null
extern "C" long long nv_check_recip_f(uint32_t* out_rec) {
  Buf B;
  if (!B.ok()) return -1;
  for (uint32_t e = 1; e <= 252; ++e) hipLaunchKernelGGL(recip_kernel, dim3(1u << 15), dim3(256), 0, 0, e, B.bad, B.rec);
  if (hipGetLastError() != hipSuccess) return -1;
  return B.finish(out_rec);
}

// every float from 2^-96 (exponent field 31) through +inf; out_rec: 33 uint32 (count, then pairs (x, s))
extern "C" long long nv_check_sqrt_rn(uint32_t* out_rec) {
  Buf B;
  if (!B.ok()) return -1;
  for (uint32_t e = 31; e <= 255; ++e) hipLaunchKernelGGL(sqrt_kernel, dim3(1u << 15), dim3(256), 0, 0, e, B.bad, B.rec);
  if (hipGetLastError() != hipSuccess) return -1;
  return B.finish(out_rec);
}

// b significands [b_lo, b_hi) (b_hi <= 2^23, multiples of 2^18 apart), every astep-th a significand;
// out_rec: 33 uint32 (count, then pairs (a, b)), may be null
extern "C" long long nv_check_div_f(uint32_t b_lo, uint32_t b_hi, uint32_t astep, uint32_t* out_rec) {
  if (b_hi > (1u << 23) || b_lo >= b_hi || (b_lo & 0x3ffffu) || (b_hi & 0x3ffffu) || astep == 0) return -1;
  Buf B;
  if (!B.ok()) return -1;
  for (uint32_t s = b_lo; s < b_hi; s += 1u << 18) {
    hipLaunchKernelGGL(pairs_kernel, dim3(1u << 10), dim3(256), 0, 0, s, astep, B.bad, B.rec);
    if (hipGetLastError() != hipSuccess) return -1;
    if (hipDeviceSynchronize() != hipSuccess) return -1;
  }
  return B.finish(out_rec);
}

// the product's division on caller-chosen operands: groups of G in {1, 3, 4} numerators per divisor; a and q hold
// groups * G floats, b holds groups floats (device pointers); synchronous, returns 0 or -1
extern "C" int nv_div_fn_apply(const float* a, const float* b, float* q, int groups, int G) {
  if (groups <= 0 || !a || !b || !q) return -1;
  const int blocks = (groups + 255) / 256;
  if (G == 1) hipLaunchKernelGGL(apply_kernel<1>, dim3(blocks), dim3(256), 0, 0, a, b, q, groups);
  else if (G == 3) hipLaunchKernelGGL(apply_kernel<3>, dim3(blocks), dim3(256), 0, 0, a, b, q, groups);
  else if (G == 4) hipLaunchKernelGGL(apply_kernel<4>, dim3(blocks), dim3(256), 0, 0, a, b, q, groups);
  else return -1;
  if (hipGetLastError() != hipSuccess) return -1;
  return hipDeviceSynchronize() == hipSuccess ? 0 : -1;
}

// s[i] = sqrt_rn(x[i]) on caller-chosen operands (device pointers); synchronous, returns 0 or -1
extern "C" int nv_sqrt_rn_apply(const float* x, float* s, int n) {
  if (n <= 0 || !x || !s) return -1;
  hipLaunchKernelGGL(sqrt_apply_kernel, dim3((n + 255) / 256), dim3(256), 0, 0, x, s, n);
  if (hipGetLastError() != hipSuccess) return -1;
  return hipDeviceSynchronize() == hipSuccess ? 0 : -1;
}
