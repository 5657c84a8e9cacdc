// hydro.hip - the fused finite-volume RK stage and the CFL signal-speed reduction for gfx950.
//
// Reference: hydro_stage_hll (sim.py:423-498), nebula3d.hydro_stage (nebula3d.py:76-151), flux_hll
// (sim.py:397-420), face_flux (nebula3d.py:54-72), limiters (nebula3d.py:40-51, sim.py:366-394),
// max_speed (nebula3d.py:154-162).
//
// Layout: structure-of-arrays fp32 fields, C-order [nx][ny][nz] (z fastest), as the reference stores them.
//
// Work decomposition. A workgroup owns a TK x TJ column of cells in (z, y) - one wavefront per y row, the 64
// lanes along the contiguous z axis - and marches along x over `ci` planes. Per plane:
//   * every thread keeps its own cell column's primitives (rho floored, u, v, w) and c^2, c for planes
//     i-1 .. i+2 in registers (a sliding window: the x stencil needs no neighbours and no halo), so the three
//     divisions of nebula3d.prim and the square root of c^2 happen once per cell;
//   * the plane's primitives go to LDS with a 2-cell y/z halo (the halo cells are the only ones converted
//     twice: (TJ+4)(TK+4) - TJ TK of them per TJ TK cells);
//   * each face flux is evaluated once: the x face i+1/2 in registers (it is the left face of plane i+1), the
//     low y and z faces of every cell from LDS into LDS face buffers, plus the TK + TJ faces on the tile's
//     high y and z boundaries; a cell then differences its two faces per axis.
// The reference evaluates six face fluxes per cell (each face twice) from 63 clamped loads; the arithmetic per
// face is identical, so the results are the same bits.
//
// Exactness: no FMA contraction (-ffp-contract=off), IEEE square root, and every division either IEEE or the
// shared-reciprocal form of nv_divide.h, which returns the identical correctly rounded quotient. Results equal
// the CPU oracle bit for bit (up to the sign of zero, which never feeds a nonzero result here).
#include "nv_common.h"
#include "nv_divide.h"

namespace nv {

constexpr int TK = 64;                      // z extent of a tile: one wavefront along the contiguous axis
constexpr int TJ = 8;                       // y rows per tile, one wavefront each
constexpr int NT = TK * TJ;                 // threads per workgroup
constexpr int PK = TK + 4, PJ = TJ + 4;     // primitive plane with a 2-cell halo
constexpr int CK = TK + 2, CJ = TJ + 2;     // c^2 / c plane with a 1-cell halo
constexpr int NHALO = 4 * PK + 4 * TJ;      // halo cells of the primitive plane
constexpr int NCHALO = 2 * CK + 2 * TJ;     // halo cells of the c plane

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

__device__ __forceinline__ float comp(const float4& q, int c) {
  return c == 0 ? q.x : (c == 1 ? q.y : (c == 2 ? q.z : q.w));
}

// Face flux between cells b and c along axis AX (0=x, 1=y, 2=z) from the primitive states of cells a, b, c, d
// and c^2, c of cells b and c. HLL (sim.py:397-420) or Rusanov (nebula3d.py:54-72, always minmod). The
// reference's one-hot axis vector e is folded in at compile time: dot(v, e) = v[AX+1], e*(r c^2) touches
// component AX+1 only.
template <int AX, int SOLVER, int LIM>
__device__ __forceinline__ float4 face_flux(const float4& qa, const float4& qb, const float4& qc, const float4& qd,
                                            float c2b, float c2c, float cb, float cc, float floor_) {
  constexpr int LM = SOLVER == NV_SOLVER_HLL ? LIM : NV_LIM_MINMOD;
  float L[4], R[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const float a = comp(qa, c), b = comp(qb, c), cv = comp(qc, c), d = comp(qd, c);
    const float d1 = b - a, d2 = cv - b, d3 = d - cv;
    L[c] = b + 0.5f * slope<LM>(d1, d2);
    R[c] = cv - 0.5f * slope<LM>(d2, d3);
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
  float F[4];
  if constexpr (SOLVER == NV_SOLVER_RUSANOV) {
    const float amax = fmaxf(fabsf(unL) + cb, fabsf(unR) + cc);
    const float ha = 0.5f * amax;
#pragma unroll
    for (int c = 0; c < 4; ++c) F[c] = 0.5f * (FL[c] + FR[c]) - ha * (UR[c] - UL[c]);
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
      const double rden = recip_d(SR - SL);
#pragma unroll
      for (int c = 0; c < 4; ++c) F[c] = div_r((SR * FL[c] - SL * FR[c]) + ss * (UR[c] - UL[c]), rden);
    }
  }
  return make_float4(F[0], F[1], F[2], F[3]);
}

// primitive state of a cell (nebula3d.prim, nebula3d.py:28-32)
__device__ __forceinline__ float4 prim(float rho, float mx, float my, float mz, float floor_) {
  const float r = fmaxf(rho, floor_);
  const double rr = recip_d(r);
  return make_float4(r, div_r(mx, rr), div_r(my, rr), div_r(mz, rr));
}

template <int SOLVER, int LIM, bool SMAX>
__global__ __launch_bounds__(NT, 2) void hydro_stage_kernel(HydroArgs A) {
  __shared__ float4 sq[PJ * PK];            // primitives of plane i, 2-cell halo
  __shared__ float2 sc[CJ * CK];            // (c^2, c) of plane i, 1-cell halo
  __shared__ float4 fy[(TJ + 1) * TK];      // y-face fluxes: row r is the face between rows r-1 and r
  __shared__ float4 fz[TJ * (TK + 1)];      // z-face fluxes: column t is the face between columns t-1 and t
  __shared__ float red[NT / 64];

  const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * TK + tx;
  const int nx = A.nx, ny = A.ny, nz = A.nz;
  const int k0 = blockIdx.x * TK, j0 = blockIdx.y * TJ, i0 = blockIdx.z * A.ci;
  const int i1 = min(i0 + A.ci, nx);
  const int k = k0 + tx, j = j0 + ty;
  const bool active = (k < nz) && (j < ny);
  const long long plane = (long long)ny * nz;
  // own column, clamped (threads past the domain edge still supply the clamped ghost values of ldc)
  const long long col = (long long)min(j, ny - 1) * nz + min(k, nz - 1);
  const float floor_ = A.floor_, cn2 = A.cn2, dci = A.ci2 - A.cn2;

  // register window of the own column: planes i-1, i, i+1, i+2
  float4 qw[4];
  float c2w[2], cw[2];          // c^2 and c of planes i, i+1
  float4 Uraw[3];               // conserved state of planes i, i+1, i+2
  float xraw[3];                // xi of planes i, i+1, i+2
  auto load_col = [&](int p, float4& U, float& x) {
    const long long o = (long long)clampi(p, 0, nx - 1) * plane + col;
    U = make_float4(A.rho[o], A.mx[o], A.my[o], A.mz[o]);
    x = A.xi[o];
  };

  // prologue: planes i0-2 .. i0+1 give the left x face of plane i0
  float4 fLx;
  {
    float4 U;
    float x;
    float4 q4[4];
    float c2p[2], cp[2];
    for (int s = 0; s < 4; ++s) {
      load_col(i0 - 2 + s, U, x);
      q4[s] = prim(U.x, U.y, U.z, U.w, floor_);
      if (s >= 1 && s <= 3) {
        const float c2 = cn2 + x * dci;
        if (s < 3) {
          c2p[s - 1] = c2;
          cp[s - 1] = sqrtf(c2);
        } else {
          c2w[1] = c2;   // plane i0+1
          cw[1] = sqrtf(c2);
        }
        if (s >= 2) {
          Uraw[s - 2] = U;
          xraw[s - 2] = x;
        }
      }
    }
    fLx = face_flux<0, SOLVER, LIM>(q4[0], q4[1], q4[2], q4[3], c2p[0], c2p[1], cp[0], cp[1], floor_);
    qw[0] = q4[1];
    qw[1] = q4[2];
    qw[2] = q4[3];
    c2w[0] = c2p[1];
    cw[0] = cp[1];
  }
  // plane i0+2 completes the window (qw[3])
  {
    load_col(i0 + 2, Uraw[2], xraw[2]);
    qw[3] = prim(Uraw[2].x, Uraw[2].y, Uraw[2].z, Uraw[2].w, floor_);
  }

  const double rdx = recip_d(A.dx);
  const float dx = A.dx;
  float tmax = 0.0f;

  for (int i = i0; i < i1; ++i) {
    // next window plane (i+3): issued now, consumed at the end of this iteration
    float4 Unext;
    float xnext;
    load_col(i + 3, Unext, xnext);
    float4 U0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const long long o = (long long)i * plane + col;
    if (!A.u0_is_u && active) U0 = make_float4(A.r0[o], A.mx0[o], A.my0[o], A.mz0[o]);

    // ---- A: plane i into LDS (own cell from the window, halo from global)
    sq[(ty + 2) * PK + tx + 2] = qw[1];
    sc[(ty + 1) * CK + tx + 1] = make_float2(c2w[0], cw[0]);
    const long long pbase = (long long)i * plane;
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
      const long long oh = pbase + (long long)clampi(j0 + jj - 2, 0, ny - 1) * nz + clampi(k0 + kk - 2, 0, nz - 1);
      sq[jj * PK + kk] = prim(A.rho[oh], A.mx[oh], A.my[oh], A.mz[oh], floor_);
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
      const long long oh = pbase + (long long)clampi(j0 + jj - 1, 0, ny - 1) * nz + clampi(k0 + kk - 1, 0, nz - 1);
      const float c2 = cn2 + A.xi[oh] * dci;
      sc[jj * CK + kk] = make_float2(c2, sqrtf(c2));
    }
    __syncthreads();

    // ---- C: face fluxes. x in registers; low y and z faces of every cell (+ tile-boundary faces) to LDS
    const float4 fRx = face_flux<0, SOLVER, LIM>(qw[0], qw[1], qw[2], qw[3], c2w[0], c2w[1], cw[0], cw[1], floor_);
    {
      const int cq = (ty + 2) * PK + tx + 2;
      const int cc = (ty + 1) * CK + tx + 1;
      const float2 cy0 = sc[cc - CK], cy1 = sc[cc];
      fy[ty * TK + tx] = face_flux<1, SOLVER, LIM>(sq[cq - 2 * PK], sq[cq - PK], sq[cq], sq[cq + PK], cy0.x, cy1.x,
                                                   cy0.y, cy1.y, floor_);
      const float2 cz0 = sc[cc - 1];
      fz[ty * (TK + 1) + tx] = face_flux<2, SOLVER, LIM>(sq[cq - 2], sq[cq - 1], sq[cq], sq[cq + 1], cz0.x, cy1.x,
                                                         cz0.y, cy1.y, floor_);
      if (tid < TK) {
        // y face between rows TJ-1 and TJ of the tile, column tid
        const int bq = (TJ + 2) * PK + tid + 2;
        const int bc = (TJ + 1) * CK + tid + 1;
        const float2 a0 = sc[bc - CK], a1 = sc[bc];
        fy[TJ * TK + tid] = face_flux<1, SOLVER, LIM>(sq[bq - 2 * PK], sq[bq - PK], sq[bq], sq[bq + PK], a0.x,
                                                      a1.x, a0.y, a1.y, floor_);
      } else if (tid < TK + TJ) {
        // z face between columns TK-1 and TK of row r
        const int r = tid - TK;
        const int bq = (r + 2) * PK + TK + 2;
        const int bc = (r + 1) * CK + TK + 1;
        const float2 a0 = sc[bc - 1], a1 = sc[bc];
        fz[r * (TK + 1) + TK] = face_flux<2, SOLVER, LIM>(sq[bq - 2], sq[bq - 1], sq[bq], sq[bq + 1], a0.x, a1.x,
                                                          a0.y, a1.y, floor_);
      }
    }
    __syncthreads();

    // ---- E: update of the own cell
    if (active) {
      const float4 fLy = fy[ty * TK + tx], fRy = fy[(ty + 1) * TK + tx];
      const float4 fLz = fz[ty * (TK + 1) + tx], fRz = fz[ty * (TK + 1) + tx + 1];
      float div[4];
#pragma unroll
      for (int c = 0; c < 4; ++c)
        div[c] = ((0.0f + (comp(fRx, c) - comp(fLx, c))) + (comp(fRy, c) - comp(fLy, c))) +
                 (comp(fRz, c) - comp(fLz, c));
      const float4 Uc = Uraw[0];
      const float U[4] = {Uc.x, Uc.y, Uc.z, Uc.w};
      float dU[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) dU[c] = div_r(-div[c], rdx);
      const float p0 = ((float)i + 0.5f) * dx, p1 = ((float)j + 0.5f) * dx, p2 = ((float)k + 0.5f) * dx;
      const float x_ion = xraw[0];
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

      const float U0v[4] = {A.u0_is_u ? U[0] : U0.x, A.u0_is_u ? U[1] : U0.y, A.u0_is_u ? U[2] : U0.z,
                            A.u0_is_u ? U[3] : U0.w};
      float out[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) out[c] = A.a * U0v[c] + A.b * (U[c] + A.dt * dU[c]);
      const float r = fmaxf(out[0], floor_);
      const double rr = recip_d(r);
      float v0 = div_r(out[1], rr), v1 = div_r(out[2], rr), v2 = div_r(out[3], rr);
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
        const float s = ((fabsf(div_r(m0, rr)) + fabsf(div_r(m1, rr))) + fabsf(div_r(m2, rr))) +
                        sqrtf(cn2 + x_ion * dci);
        tmax = fmaxf(tmax, s);
      }
    }
    fLx = fRx;

    // ---- shift the window by one plane
    qw[0] = qw[1];
    qw[1] = qw[2];
    qw[2] = qw[3];
    qw[3] = prim(Unext.x, Unext.y, Unext.z, Unext.w, floor_);
    c2w[0] = c2w[1];
    cw[0] = cw[1];
    c2w[1] = cn2 + xraw[2] * dci;   // plane i+2 becomes plane (i+1)+1
    cw[1] = sqrtf(c2w[1]);
    Uraw[0] = Uraw[1];
    Uraw[1] = Uraw[2];
    Uraw[2] = Unext;
    xraw[0] = xraw[1];
    xraw[1] = xraw[2];
    xraw[2] = xnext;
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
  NV_REQUIRE(floor_ > 0.0f && dx > 0.0f, "nv_hydro_stage: floor and dx must be > 0");
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
