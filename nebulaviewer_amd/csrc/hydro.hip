// hydro.hip - the fused finite-volume RK stage and the CFL signal-speed reduction for gfx950.
//
// Reference: hydro_stage_hll (sim.py:423-498), nebula3d.hydro_stage (nebula3d.py:76-151), flux_hll
// (sim.py:397-420), face_flux (nebula3d.py:54-72), limiters (nebula3d.py:40-51, sim.py:366-394),
// max_speed (nebula3d.py:154-162).
//
// Layout: structure-of-arrays fp32 fields, C-order [nx][ny][nz] (z fastest), as the reference stores them.
// Tiling: a workgroup owns a TK x TJ column of cells in (z, y) and marches along x over `ci` planes. Each
// x-plane is loaded once (with a 2-cell halo in y and z), converted to primitives (the 3 divisions of
// nebula3d.prim happen once per loaded cell, not 15 times per cell as in the reference stencil) and kept in
// an LDS ring of 5 planes; the x-direction stencil reads the ring, y and z read the current plane. The x-face
// flux is computed once per face (the right face of plane i is the left face of plane i+1, carried in
// registers); the y/z faces are evaluated per cell.
//
// Arithmetic follows the reference expression by expression with no FMA contraction (-ffp-contract=off) and
// IEEE division/sqrt, so results equal the CPU oracle bit for bit (up to the sign of zero, which never feeds
// a nonzero result: every divisor here is strictly positive).
#include "nv_common.h"

namespace nv {

constexpr int TK = 64;              // z extent of a tile: one wavefront along the contiguous axis
constexpr int TJ = 4;               // y rows of a tile (one wavefront each)
constexpr int NT = TK * TJ;         // threads per workgroup
constexpr int PK = TK + 4, PJ = TJ + 4, PN = PK * PJ;   // primitive plane with a 2-cell halo
constexpr int CK = TK + 2, CJ = TJ + 2, CN = CK * CJ;   // c^2 plane with a 1-cell halo
constexpr int NQ = 5;               // primitive plane ring (planes i-2 .. i+2 live)
constexpr int NC = 3;               // c^2 plane ring (planes i, i+1 read; i+1 written)

struct HydroArgs {
  const float *rho, *mx, *my, *mz, *xi, *r0, *mx0, *my0, *mz0;
  float *o_r, *o_mx, *o_my, *o_mz;
  const float4 *stars, *jets;
  float* smax;
  int nx, ny, nz, nstars, ci, u0_is_u;
  float wind, jet_cos, jet_r, pulse, dt, dx, a, b, floor_, cn2, ci2, vmax;
};

// slope limiters: minmod (nebula3d.py:40-46), van Leer (sim.py:375-380), MC (sim.py:366-372)
template <int LIM>
__device__ __forceinline__ float slope(float a, float b) {
  const float p = a * b;
  if constexpr (LIM == NV_LIM_VANLEER) {
    const float v = 2.0f * a * b / (a + b);
    return p <= 0.0f ? 0.0f : v;
  } else if constexpr (LIM == NV_LIM_MC) {
    const float s = a > 0.0f ? 1.0f : -1.0f;
    const float v = s * fminf(fminf(2.0f * fabsf(a), 2.0f * fabsf(b)), 0.5f * fabsf(a + b));
    return p <= 0.0f ? 0.0f : v;
  } else {
    const float m = fabsf(a) < fabsf(b) ? a : b;
    return p <= 0.0f ? 0.0f : m;
  }
}

// Face flux between cells b and c along axis AX (0=x, 1=y, 2=z) from the primitive states of cells a, b, c,
// d. HLL (sim.py:397-420) or Rusanov (nebula3d.py:54-72; always minmod). The one-hot axis vector e of the
// reference is folded in at compile time: dot(v, e) = v[AX+1], e * (r c^2) touches component AX+1 only.
template <int AX, int SOLVER, int LIM>
__device__ __forceinline__ void face_flux(const float* qa, const float* qb, const float* qc, const float* qd,
                                          float c2b, float c2c, float floor_, float* F) {
  constexpr int LM = SOLVER == NV_SOLVER_HLL ? LIM : NV_LIM_MINMOD;
  float L[4], R[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    L[c] = qb[c] + 0.5f * slope<LM>(qb[c] - qa[c], qc[c] - qb[c]);
    R[c] = qc[c] - 0.5f * slope<LM>(qc[c] - qb[c], qd[c] - qc[c]);
  }
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
  FL[AX + 1] = FL[AX + 1] + rL * c2b;
  FR[AX + 1] = FR[AX + 1] + rR * c2c;
  if constexpr (SOLVER == NV_SOLVER_RUSANOV) {
    const float amax = fmaxf(fabsf(unL) + sqrtf(c2b), fabsf(unR) + sqrtf(c2c));
    const float ha = 0.5f * amax;
#pragma unroll
    for (int c = 0; c < 4; ++c) F[c] = 0.5f * (FL[c] + FR[c]) - ha * (UR[c] - UL[c]);
  } else {
    const float cb = sqrtf(c2b), cc = sqrtf(c2c);
    const float SL = fminf(unL - cb, unR - cc);
    const float SR = fmaxf(unL + cb, unR + cc);
    if (SL >= 0.0f) {
#pragma unroll
      for (int c = 0; c < 4; ++c) F[c] = FL[c];
    } else if (SR <= 0.0f) {
#pragma unroll
      for (int c = 0; c < 4; ++c) F[c] = FR[c];
    } else {
      const float ss = SL * SR, den = SR - SL;
#pragma unroll
      for (int c = 0; c < 4; ++c) F[c] = ((SR * FL[c] - SL * FR[c]) + ss * (UR[c] - UL[c])) / den;
    }
  }
}

__device__ __forceinline__ int ring(int p, int n) { return ((p % n) + n) % n; }

template <int SOLVER, int LIM, bool SMAX>
__global__ __launch_bounds__(NT) void hydro_stage_kernel(HydroArgs A) {
  __shared__ float sq[NQ][4][PN];
  __shared__ float sc[NC][CN];
  __shared__ float red[NT / 64];

  const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * TK + tx;
  const int nx = A.nx, ny = A.ny, nz = A.nz;
  const int k0 = blockIdx.x * TK, j0 = blockIdx.y * TJ, i0 = blockIdx.z * A.ci;
  const int i1 = min(i0 + A.ci, nx);
  const int k = k0 + tx, j = j0 + ty;
  const bool active = (k < nz) && (j < ny);
  const float floor_ = A.floor_, cn2 = A.cn2, ci2 = A.ci2, dci = ci2 - cn2;

  // primitive plane p (clamped as ldc, nebula3d.py:23-25) into its ring slot
  auto load_q = [&](int p) {
    const int s = ring(p, NQ);
    const long long base = (long long)clampi(p, 0, nx - 1) * ny * nz;
    for (int t = tid; t < PN; t += NT) {
      const int jj = t / PK, kk = t - jj * PK;
      const long long o = base + (long long)clampi(j0 + jj - 2, 0, ny - 1) * nz + clampi(k0 + kk - 2, 0, nz - 1);
      const float r = fmaxf(A.rho[o], floor_);
      sq[s][0][t] = r;
      sq[s][1][t] = A.mx[o] / r;
      sq[s][2][t] = A.my[o] / r;
      sq[s][3][t] = A.mz[o] / r;
    }
  };
  // c^2 plane (csq, nebula3d.py:35-37)
  auto load_c = [&](int p) {
    const int s = ring(p, NC);
    const long long base = (long long)clampi(p, 0, nx - 1) * ny * nz;
    for (int t = tid; t < CN; t += NT) {
      const int jj = t / CK, kk = t - jj * CK;
      const long long o = base + (long long)clampi(j0 + jj - 1, 0, ny - 1) * nz + clampi(k0 + kk - 1, 0, nz - 1);
      sc[s][t] = cn2 + A.xi[o] * dci;
    }
  };
  auto getq = [&](int p, int dj, int dk, float* q) {
    const int s = ring(p, NQ), t = (ty + 2 + dj) * PK + tx + 2 + dk;
#pragma unroll
    for (int c = 0; c < 4; ++c) q[c] = sq[s][c][t];
  };
  auto getc = [&](int p, int dj, int dk) { return sc[ring(p, NC)][(ty + 1 + dj) * CK + tx + 1 + dk]; };

  load_q(i0 - 2);
  load_q(i0 - 1);
  load_q(i0);
  load_q(i0 + 1);
  load_c(i0 - 1);
  load_c(i0);
  __syncthreads();

  float fLx[4];
  {
    float qa[4], qb[4], qc[4], qd[4];
    getq(i0 - 2, 0, 0, qa);
    getq(i0 - 1, 0, 0, qb);
    getq(i0, 0, 0, qc);
    getq(i0 + 1, 0, 0, qd);
    face_flux<0, SOLVER, LIM>(qa, qb, qc, qd, getc(i0 - 1, 0, 0), getc(i0, 0, 0), floor_, fLx);
  }

  float tmax = 0.0f;
  const float dx = A.dx;
  for (int i = i0; i < i1; ++i) {
    load_q(i + 2);
    load_c(i + 1);
    __syncthreads();
    if (active) {
      float div[4], fR[4], fL[4];
      float q[5][4];
      // x faces: right face from the ring, left face carried from the previous plane
      getq(i - 1, 0, 0, q[0]);
      getq(i, 0, 0, q[1]);
      getq(i + 1, 0, 0, q[2]);
      getq(i + 2, 0, 0, q[3]);
      const float c0 = getc(i, 0, 0);
      face_flux<0, SOLVER, LIM>(q[0], q[1], q[2], q[3], c0, getc(i + 1, 0, 0), floor_, fR);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        div[c] = 0.0f + (fR[c] - fLx[c]);
        fLx[c] = fR[c];
      }
      // y faces
#pragma unroll
      for (int d = 0; d < 5; ++d) getq(i, d - 2, 0, q[d]);
      face_flux<1, SOLVER, LIM>(q[1], q[2], q[3], q[4], c0, getc(i, 1, 0), floor_, fR);
      face_flux<1, SOLVER, LIM>(q[0], q[1], q[2], q[3], getc(i, -1, 0), c0, floor_, fL);
#pragma unroll
      for (int c = 0; c < 4; ++c) div[c] = div[c] + (fR[c] - fL[c]);
      // z faces
#pragma unroll
      for (int d = 0; d < 5; ++d) getq(i, 0, d - 2, q[d]);
      face_flux<2, SOLVER, LIM>(q[1], q[2], q[3], q[4], c0, getc(i, 0, 1), floor_, fR);
      face_flux<2, SOLVER, LIM>(q[0], q[1], q[2], q[3], getc(i, 0, -1), c0, floor_, fL);
#pragma unroll
      for (int c = 0; c < 4; ++c) div[c] = div[c] + (fR[c] - fL[c]);

      const long long o = cell(i, j, k, ny, nz);
      const float U[4] = {A.rho[o], A.mx[o], A.my[o], A.mz[o]};
      float dU[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) dU[c] = -div[c] / dx;
      const float p0 = ((float)i + 0.5f) * dx, p1 = ((float)j + 0.5f) * dx, p2 = ((float)k + 0.5f) * dx;
      const float x_ion = A.xi[o];
      if (A.wind > 0.0f && x_ion > 0.01f) {
        // radiation pressure / wind on ionized gas (sim.py:463-473)
        float acc0 = 0.0f, acc1 = 0.0f, acc2 = 0.0f;
        for (int s = 0; s < A.nstars; ++s) {
          const float4 st = A.stars[s];
          if (st.w > 0.0f) {
            const float r0 = p0 - st.x, r1 = p1 - st.y, r2v = p2 - st.z;
            const float r2 = ((r0 * r0 + r1 * r1) + r2v * r2v) + 0.0004f;
            const float f = st.w / (r2 * sqrtf(r2));
            acc0 = acc0 + r0 * f;
            acc1 = acc1 + r1 * f;
            acc2 = acc2 + r2v * f;
          }
        }
        const float wx = A.wind * x_ion;
        dU[1] = dU[1] + U[0] * (acc0 * wx);
        dU[2] = dU[2] + U[0] * (acc1 * wx);
        dU[3] = dU[3] + U[0] * (acc2 * wx);
      }
      // bipolar jets (sim.py:474-485)
      float aj0 = 0.0f, aj1 = 0.0f, aj2 = 0.0f;
      for (int s = 0; s < A.nstars; ++s) {
        const float4 jt = A.jets[s];
        const float4 st = A.stars[s];
        if (jt.w > 0.0f && st.w > 0.0f) {
          const float r0 = p0 - st.x, r1 = p1 - st.y, r2v = p2 - st.z;
          const float r = sqrtf((r0 * r0 + r1 * r1) + r2v * r2v);
          if (r < A.jet_r && r > 1e-6f) {
            const float cth = ((r0 * jt.x + r1 * jt.y) + r2v * jt.z) / r;
            if (fabsf(cth) > A.jet_cos) {
              const float f = (cth > 0.0f ? 1.0f : -1.0f) * jt.w * A.pulse / (r * r + 0.0004f);
              aj0 = aj0 + jt.x * f;
              aj1 = aj1 + jt.y * f;
              aj2 = aj2 + jt.z * f;
            }
          }
        }
      }
      dU[1] = dU[1] + U[0] * aj0;
      dU[2] = dU[2] + U[0] * aj1;
      dU[3] = dU[3] + U[0] * aj2;

      float U0[4];
      if (A.u0_is_u) {
#pragma unroll
        for (int c = 0; c < 4; ++c) U0[c] = U[c];
      } else {
        U0[0] = A.r0[o];
        U0[1] = A.mx0[o];
        U0[2] = A.my0[o];
        U0[3] = A.mz0[o];
      }
      float out[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) out[c] = A.a * U0[c] + A.b * (U[c] + A.dt * dU[c]);
      const float r = fmaxf(out[0], floor_);
      float v0 = out[1] / r, v1 = out[2] / r, v2 = out[3] / r;
      const float sp = sqrtf((v0 * v0 + v1 * v1) + v2 * v2);
      if (sp > A.vmax) {
        const float f = A.vmax / sp;
        v0 = v0 * f;
        v1 = v1 * f;
        v2 = v2 * f;
      }
      const float m0 = r * v0, m1 = r * v1, m2 = r * v2;
      A.o_r[o] = r;
      A.o_mx[o] = m0;
      A.o_my[o] = m1;
      A.o_mz[o] = m2;
      if constexpr (SMAX) {
        // max_speed of the written state (nebula3d.py:154-162): raw stored rho, L1 norm of v
        const float s = ((fabsf(m0 / r) + fabsf(m1 / r)) + fabsf(m2 / r)) + sqrtf(cn2 + x_ion * dci);
        tmax = fmaxf(tmax, s);
      }
    }
  }
  if constexpr (SMAX) block_max_atomic(tmax, red, A.smax);
}

// max_speed (nebula3d.py:154-162)
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

using HydroKernel = void (*)(HydroArgs);

template <bool SMAX>
static HydroKernel pick_hydro(int solver, int lim) {
  if (solver == NV_SOLVER_RUSANOV) return hydro_stage_kernel<NV_SOLVER_RUSANOV, NV_LIM_MINMOD, SMAX>;
  if (lim == NV_LIM_VANLEER) return hydro_stage_kernel<NV_SOLVER_HLL, NV_LIM_VANLEER, SMAX>;
  if (lim == NV_LIM_MC) return hydro_stage_kernel<NV_SOLVER_HLL, NV_LIM_MC, SMAX>;
  return hydro_stage_kernel<NV_SOLVER_HLL, NV_LIM_MINMOD, SMAX>;
}

}  // namespace nv

extern "C" int nv_hydro_stage(const float* rho, const float* mx, const float* my, const float* mz, const float* xi,
                              const float* r0, const float* mx0, const float* my0, const float* mz0, float* o_r,
                              float* o_mx, float* o_my, float* o_mz, int nx, int ny, int nz, const float* stars,
                              int nstars, float wind, const float* jets, float jet_cos, float jet_r, float pulse,
                              float dt, float dx, float a, float b, float floor_, float cn2, float ci2, float vmax,
                              int solver, int lim, float* smax_out, void* stream) {
  using namespace nv;
  clear_error();
  NV_REQUIRE(rho && mx && my && mz && xi && r0 && mx0 && my0 && mz0 && o_r && o_mx && o_my && o_mz,
             "nv_hydro_stage: null field pointer");
  NV_REQUIRE(nx > 0 && ny > 0 && nz > 0, "nv_hydro_stage: bad grid %dx%dx%d", nx, ny, nz);
  NV_REQUIRE(nstars >= 0 && (nstars == 0 || (stars && jets)), "nv_hydro_stage: nstars=%d needs star/jet tables",
             nstars);
  NV_REQUIRE(solver == NV_SOLVER_RUSANOV || solver == NV_SOLVER_HLL, "nv_hydro_stage: bad solver %d", solver);
  NV_REQUIRE(lim >= NV_LIM_MINMOD && lim <= NV_LIM_MC, "nv_hydro_stage: bad limiter %d", lim);
  NV_REQUIRE(o_r != rho && o_mx != mx && o_my != my && o_mz != mz, "nv_hydro_stage: output aliases the input U");
  HydroArgs A;
  A.rho = rho; A.mx = mx; A.my = my; A.mz = mz; A.xi = xi;
  A.r0 = r0; A.mx0 = mx0; A.my0 = my0; A.mz0 = mz0;
  A.o_r = o_r; A.o_mx = o_mx; A.o_my = o_my; A.o_mz = o_mz;
  A.stars = reinterpret_cast<const float4*>(stars);
  A.jets = reinterpret_cast<const float4*>(jets);
  A.smax = smax_out;
  A.nx = nx; A.ny = ny; A.nz = nz; A.nstars = nstars;
  A.u0_is_u = (r0 == rho && mx0 == mx && my0 == my && mz0 == mz) ? 1 : 0;
  A.wind = wind; A.jet_cos = jet_cos; A.jet_r = jet_r; A.pulse = pulse; A.dt = dt; A.dx = dx;
  A.a = a; A.b = b; A.floor_ = floor_; A.cn2 = cn2; A.ci2 = ci2; A.vmax = vmax;
  // march depth: long enough to amortise the 4-plane prologue, short enough to fill the chip
  const long long tiles_jk = (long long)((nz + TK - 1) / TK) * ((ny + TJ - 1) / TJ);
  int ci = (int)std::min<long long>(64, std::max<long long>(4, (long long)nx * tiles_jk / 2048));
  A.ci = ci;
  dim3 grid((nz + TK - 1) / TK, (ny + TJ - 1) / TJ, (nx + ci - 1) / ci);
  dim3 block(TK, TJ, 1);
  HydroKernel kern = smax_out ? pick_hydro<true>(solver, lim) : pick_hydro<false>(solver, lim);
  hipLaunchKernelGGL(kern, grid, block, 0, as_stream(stream), A);
  return check_launch("nv_hydro_stage");
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
