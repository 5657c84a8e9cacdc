// Exhaustive checks of the float-only exact division of nebulaviewer_amd/csrc/nv_divide.h on the GPU it runs on
// (test helper, built into tests/csrc/libnv_divcheck.so):
//   nv_check_recip_f:  recip_f(b).y == RN(1/b) for every normal b below 2^126 (exponent fields 1..252);
//   nv_check_div_f:    div_fn == the IEEE quotient for all 2^46 pairs of significands a, b in [1, 2) (every
//                      astep-th a), the reference being the double-reciprocal quotient, which is correctly
//                      rounded for normal quotients.
// Both return the number of mismatches (or -1 on a HIP error) and record up to 16 offending operand pairs.
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
