/* nebula_hip.h - C ABI of the MI355X (gfx950) simulation-step kernels of nebulaviewer_amd.
 *
 * Every entry point:
 *   - takes device pointers owned by the caller (plain float / uint16_t buffers, C-order [nx][ny][nz] with
 *     z fastest, the layout of sim.py:721-733), grid sizes as ints and a hipStream_t passed as void*;
 *   - launches asynchronously on that stream and returns NV_OK (0) or a nonzero code; nv_last_error()
 *     then holds a message for the calling thread. Nothing throws across this boundary;
 *   - is stateless (no allocations, no globals written) and reentrant per stream; the one exception is
 *     nv_cast_rays_noalloc, which stages its data in a static device array of the library (see there).
 * Scalars follow the reference launch semantics: the caller computes them in double (Python floats) and
 * passes them rounded to float, exactly as a Warp launch rounds its float arguments.
 *
 * Star tables: stars = float4[nstars] of (x, y, z, w) with w = Q_eff / q_main, 0 for a star not yet on
 * (sim.py:914-925); jets = float4[nstars] of (axis x, y, z, strength).
 */
#ifndef NEBULA_HIP_H
#define NEBULA_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NV_OK 0
#define NV_ERR_ARG 1001    /* invalid argument (null pointer, bad size, bad enum) */
#define NV_ERR_LAUNCH 1002 /* the kernel launch failed (message in nv_last_error) */

#define NV_SOLVER_RUSANOV 0 /* nebula3d.face_flux (nebula3d.py:54-72) */
#define NV_SOLVER_HLL 1     /* sim.flux_hll (sim.py:397-420) */
#define NV_LIM_MINMOD 0     /* sim.py:1022 limiter codes */
#define NV_LIM_VANLEER 1
#define NV_LIM_MC 2

#define NV_PREC_F32 0      /* nv_hydro_stage */
#define NV_PREC_F32_FAST 1 /* nv_hydro_stage_fast */
#define NV_PREC_F16V 2     /* nv_hydro_stage_h (float32 rho, binary16 velocity) */

#define NV_TILE_AUTO 0     /* the library's choice from the z extent */
#define NV_TILE_64x8 1     /* 64 z x 8 y cells per workgroup (one wavefront per y row) */
#define NV_TILE_32x16 2    /* 32 z x 16 y cells per workgroup (one wavefront per two rows) */
#define NV_TILE_64x16 3    /* 64 z x 16 y cells per workgroup (1024 threads, one workgroup per CU) */
/* A tile code may also carry the march depth: NV_TILE_DEPTH(n) | shape makes each workgroup advance its column
 * over n x planes (at most the grid's nx; n >= nx marches the whole column in one workgroup); depth 0 is the
 * library's choice (up to 64 planes, enough workgroups to fill the device). Like the shape it changes the work
 * decomposition only, never a bit of the result. */
#define NV_TILE_DEPTH(n) ((n) << 8)

/* One grid of a batched RK stage: the arguments of nv_hydro_stage, with the four fields of U, U0 and the output
 * as arrays (float32, or float32 rho + binary16 velocity for NV_PREC_F16V). dt_dev: if non-null, the stage reads
 * its dt from this device float when it runs (nv_cfl_dt writes it), and dt is ignored. */
typedef struct NvStageDesc {
  const void* u[4];
  const float* xi;
  const void* u0[4];
  void* out[4];
  const float* stars;
  const float* jets;
  float* smax_out;
  const float* dt_dev;
  int nx, ny, nz, nstars;
  float wind, jet_cos, jet_r, pulse, dt, dx, a, b, floor_, cn2, ci2, vmax;
} NvStageDesc;

/* One ray-cast job of a batch: the arguments of nv_cast_rays (wide != 0: those of nv_cast_rays_wide). bricks:
 * null, or a 16-byte aligned scratch of nv_ray_bricks_size(nx, ny, nz) bytes that nv_cast_rays_batch fills with
 * rho's brick array (as nv_ray_bricks) before the casts sample it; jobs on one grid pass the same rho and bricks,
 * built once. dirs: null, or the direction table nv_ray_dirs wrote for this job's (nu, nphi), which jobs with the
 * same ray grid share (the same bits as the directions the cast computes itself). */
typedef struct NvRayDesc {
  const float* rho;
  int nx, ny, nz;
  float sx, sy, sz, Q, dr;
  int nsteps;
  float dx, rcap, bx, by, bz;
  float* rfront;
  int nu, nphi, wide;
  float* bricks;
  const float* dirs; /* null, or nv_ray_dirs(nu, nphi): the rays' directions, read instead of computed per ray */
  float* rf_range;   /* null, or nv_ray_range_floats(nu, nphi) floats whose first two nv_cast_rays_batch sets to the
                      * least and greatest entry of rfront (the rest is its scratch) */
} NvRayDesc;

/* One grid of a batched ionize pass: the arguments of nv_ionize_apply (state = rho, mx, my, mz, or rho, vx, vy, vz
 * binary16 for nv_ionize_apply_h), and optionally the stars' front ranges. */
typedef struct NvIonDesc {
  const float* rfronts;
  int nu, nphi;
  const float* star_xyz;
  int nstars;
  float dx, width;
  float* xi;
  uint16_t* dfront_h;
  float* dfront_f;
  int nx, ny, nz;
  const void* state[4];
  float cn2, ci2;
  float* smax_out;
  /* null, or per star the nv_ray_range_floats(nu, nphi) floats of its job's NvRayDesc.rf_range, one star after the
   * other: the pass then skips a star where a proven bound shows it changes neither xi nor dfront (the same bits;
   * no reference counterpart) */
  const float* rf_ranges;
} NvIonDesc;

/* Render packing parameters: the reference's PackParams struct (render.py:77-88), same member order. */
typedef struct NvPackParams {
  float c_ha[3], c_o3[3], c_s2[3], c_dg[3], c_sh[3];
  float gs, ambient, edge;
  float inv_L[3];
  float dx;
} NvPackParams;

/* Message of the last failed call on this thread ("" if none). */
const char* nv_last_error(void);
/* ABI version of this library (bumped on any signature change). */
int nv_abi_version(void);

/* One RK stage of the fused finite-volume update: MUSCL reconstruction, face fluxes on x/y/z, wind and jet
 * sources, out = a*U0 + b*(U + dt*dU), density floor, |v| <= vmax.
 * Replaces the launches of hydro_stage_hll (sim.py:423-498, solver=HLL) and nebula3d.hydro_stage
 * (nebula3d.py:76-151, solver=Rusanov) made by Sim._hydro (sim.py:1014-1028).
 * In-place use as in sim.py:1026-1028 is allowed: out may alias U0 (each cell reads its own U0 before
 * writing it); out must not alias U. U0 may alias U (stage 1).
 * smax_out: if non-null, the kernel also max-reduces |u|+|v|+|w|+c of the cells it writes into smax_out[0]
 * (float bits, atomic; the caller zeroes it), which equals nv_max_speed on the output fields and xi. */
int nv_hydro_stage(const float* rho, const float* mx, const float* my, const float* mz, const float* xi,
                   const float* r0, const float* mx0, const float* my0, const float* mz0, float* o_r,
                   float* o_mx, float* o_my, float* o_mz, int nx, int ny, int nz, const float* stars,
                   int nstars, float wind, const float* jets, float jet_cos, float jet_r, float pulse, float dt,
                   float dx, float a, float b, float floor_, float cn2, float ci2, float vmax, int solver,
                   int lim, float* smax_out, void* stream);

/* The same stage with fused multiply-adds, hardware reciprocals (one Newton step) and hardware square roots:
 * each operation within about one ulp instead of correctly rounded. Same arguments as nv_hydro_stage. */
int nv_hydro_stage_fast(const float* rho, const float* mx, const float* my, const float* mz, const float* xi,
                        const float* r0, const float* mx0, const float* my0, const float* mz0, float* o_r,
                        float* o_mx, float* o_my, float* o_mz, int nx, int ny, int nz, const float* stars,
                        int nstars, float wind, const float* jets, float jet_cos, float jet_r, float pulse, float dt,
                        float dx, float a, float b, float floor_, float cn2, float ci2, float vmax, int solver,
                        int lim, float* smax_out, void* stream);

/* The stage on the reduced-storage state: float32 rho with binary16 velocity (vx, vy, vz) for U, U0 and the
 * output (velocity, not momentum; arithmetic in float32 with the fast operations above). Replaces the same
 * reference launch as nv_hydro_stage with a 58 B/cell-step instead of an 88 B/cell-step state. smax_out
 * reduces |v| + c of the stored (rounded) state. */
int nv_hydro_stage_h(const float* rho, const uint16_t* vx, const uint16_t* vy, const uint16_t* vz,
                     const float* xi, const float* r0, const uint16_t* vx0, const uint16_t* vy0,
                     const uint16_t* vz0, float* o_r, uint16_t* o_vx, uint16_t* o_vy, uint16_t* o_vz, int nx,
                     int ny, int nz, const float* stars, int nstars, float wind, const float* jets, float jet_cos,
                     float jet_r, float pulse, float dt, float dx, float a, float b, float floor_, float cn2,
                     float ci2, float vmax, int solver, int lim, float* smax_out, void* stream);

/* nv_hydro_stage (or _fast / _h, by precision) from a descriptor, with an explicit workgroup tile shape
 * (NV_TILE_*): the shape changes the work decomposition only, never a bit of the result. The entry point an
 * autotuner uses; nv_hydro_stage and its variants take NV_TILE_AUTO. Replaces the same launch as nv_hydro_stage. */
int nv_hydro_stage_ex(const NvStageDesc* d, int precision, int solver, int lim, int tile, void* stream);

/* Batched RK stage: many independent grids (uneven sizes) advance in one launch over the concatenated tile
 * list. No reference counterpart (the reference steps one grid per Sim, sim.py:1014-1028); each member's result
 * is bit-identical to nv_hydro_stage (or _fast / _h, by precision) called on that member alone.
 * Two steps, so that a launch needs nothing from the host but a pointer (hipGraph-capturable):
 *   nv_hydro_stage_batch_plan: validates the n members (as nv_hydro_stage does, plus: all members set smax_out or
 *     none; solver, limiter, precision and the tile shape, NV_TILE_AUTO for the library's choice, are shared) and
 *     writes a plan of nv_hydro_stage_batch_plan_size(n) bytes into host memory `plan`;
 *   nv_hydro_stage_batch: launches the stage from the plan: plan_host (the host blob) and plan_dev (a device
 *     copy of the same bytes, 8-byte aligned, which must stay alive and unchanged while launches use it). */
long long nv_hydro_stage_batch_plan_size(int n);
int nv_hydro_stage_batch_plan(const NvStageDesc* members, int n, int precision, int solver, int lim, int tile,
                              void* plan, long long plan_bytes);
int nv_hydro_stage_batch(const void* plan_host, const void* plan_dev, void* stream);
/* The same stage from the same plan, as a persistent launch: as many workgroups as the device holds at once, fed by
 * a tile work queue in the plan's device copy (per-XCD runs of tiles, claimed by atomic adds, stolen across XCDs
 * when a run is exhausted; the last workgroup resets the counters). Bit-identical to nv_hydro_stage_batch; two
 * launches of one plan must not overlap (one stream). No reference counterpart (the reference has no batching). */
int nv_hydro_stage_batch_queue(const void* plan_host, const void* plan_dev, void* stream);
/* The persistent launch with its residency given: wgs_per_cu workgroups per CU (the grid's waves per CU are
 * wgs_per_cu times the tile's wavefronts), capped at what the kernel's registers and LDS allow; 0 for that maximum,
 * which is nv_hydro_stage_batch_queue. Bit-identical to nv_hydro_stage_batch at every residency. A launch choice for
 * the batch tuner; no reference counterpart. */
int nv_hydro_stage_batch_queue_ex(const void* plan_host, const void* plan_dev, int wgs_per_cu, void* stream);

/* Mass budget in FP64 on the GPU (the north star's global conservation sums): out[0] += sum of rho over the grid,
 * out[1] += the net outward mass flux through the six domain faces, each face's float flux (the one the exact
 * stage adds to the boundary cell: with clamped ghosts both reconstructed states are the cell's own) summed in
 * double. For a stage out = a U0 + b (U + dt dU) that the density floor does not touch,
 * sum rho_out = a sum rho0 + b (sum rho - (dt / dx) out[1]) up to the rounding of each cell's update. The caller
 * zeroes out[0..1]. Replaces the d_bflux accumulation a conservation check needs; no single reference line. */
int nv_conservation_sums(const float* rho, const float* mx, const float* my, const float* mz, const float* xi,
                         int nx, int ny, int nz, float floor_, float cn2, float ci2, int solver, double* out,
                         void* stream);

/* Device-side timestep of n grids, so a step needs no host round trip: for each s,
 *   dt = cfl_dx[s] / (double)smax[s]  (cfl_dx = cfl * dx computed on the host in double, as sim.py:1011 does),
 *   dt = min(dt, dt_cap[s]) if dt_cap is non-null and dt_cap[s] > 0,
 *   dt_f[s] = (float)dt (what the stages read through NvStageDesc.dt_dev), dt_d[s] = dt and t[s] += dt when those
 *   are non-null (the clock of sim.py:1111), and smax[s] = 0 (ready for the next reduction).
 * Replaces the host side of Sim._compute_dt (sim.py:1009-1012) and its readback. */
int nv_cfl_dt(float* smax, const double* cfl_dx, const double* dt_cap, float* dt_f, double* dt_d, double* t, int n,
              void* stream);

/* CFL signal speed: out[0] = max(out[0], max over cells of |u|+|v|+|w| + sqrt(cn2 + xi (ci2 - cn2))), raw rho.
 * Replaces nebula3d.max_speed (nebula3d.py:154-162) as launched by Sim._compute_dt (sim.py:1009-1012). */
int nv_max_speed(const float* rho, const float* mx, const float* my, const float* mz, const float* xi, int nx,
                 int ny, int nz, float cn2, float ci2, float* out, void* stream);

/* nv_max_speed for the binary16-velocity state: max of |vx|+|vy|+|vz| + c (max_speed, nebula3d.py:154-162). */
int nv_max_speed_h(const uint16_t* vx, const uint16_t* vy, const uint16_t* vz, const float* xi, int nx, int ny,
                   int nz, float cn2, float ci2, float* out, void* stream);

/* Stromgren front radius per direction: rfront[nu][nphi]. Replaces nebula3d.cast_rays (nebula3d.py:247-270),
 * launched per star by Sim.ionize (sim.py:936-937). */
int nv_cast_rays(const float* rho, int nx, int ny, int nz, float sx, float sy, float sz, float Q, float dr,
                 int nsteps, float dx, float rcap, float bx, float by, float bz, float* rfront, int nu, int nphi,
                 void* stream);

/* nv_cast_rays with the recombination column of long rays in FP64: a ray whose marched length exceeds 512 cells
 * accumulates the column, tests col + inc > Q and interpolates the front radius in double (the radius rounded
 * to float at the end); shorter rays give nv_cast_rays' result bit for bit. No reference counterpart: the
 * reference sums in float (nebula3d.py:262-268), whose threshold test drifts on long rays. */
int nv_cast_rays_wide(const float* rho, int nx, int ny, int nz, float sx, float sy, float sz, float Q, float dr,
                      int nsteps, float dx, float rcap, float bx, float by, float bz, float* rfront, int nu,
                      int nphi, void* stream);

/* The brick array of a density field for nv_cast_rays_ex, with one ghost layer: for p in [0, nx+2), q in
 * [0, ny+1), r in [0, nz+1), bricks[p][q][r] = (rho(i, ja, ka), rho(i, ja, kb), rho(i, jb, ka), rho(i, jb, kb)) as
 * 4 floats, i = clamp(p-1), ja = clamp(q-1), jb = clamp(q), ka = clamp(r-1), kb = clamp(r) (clamped to the grid):
 * the eight corners of any trilinear sample are the bricks at (i0+1, j0+1, k0+1) and (i0+2, j0+1, k0+1), with no
 * clamping in the sampler. nv_ray_bricks_size(nx, ny, nz) bytes, 16-byte aligned (about 16 B per cell: Sim's RK
 * stage block, which is free while ionize runs, is allocated that large). Built once per ionization. No reference
 * counterpart: it is the sampler's data layout, not a reference step. */
long long nv_ray_bricks_size(int nx, int ny, int nz);

/* The direction of every ray of an (nu, nphi) grid (nebula3d.py:252-257): dirs[4 (iu nphi + ip) + c] = d_c for
 * c < 3 (the fourth float 0), a 16-byte aligned buffer of 16 nu nphi bytes. No reference counterpart (the reference
 * evaluates the direction in every ray's thread, nebula3d.py:247-257); a batch's casts read it (NvRayDesc.dirs). */
int nv_ray_dirs(int nu, int nphi, float* dirs, void* stream);

/* Floats of a ray job's front range (NvRayDesc.rf_range) for an (nu, nphi) ray grid: the (least, greatest) pair and
 * the casts' partial ranges; -1 for a bad grid. No reference counterpart (the ranges only let the batched ionize
 * pass skip work, NvIonDesc.rf_ranges). */
long long nv_ray_range_floats(int nu, int nphi);
int nv_ray_bricks(const float* rho, int nx, int ny, int nz, float* bricks, void* stream);

/* nv_cast_rays (wide == 0) or nv_cast_rays_wide (wide != 0), sampling the density from its brick array when
 * bricks is non-null: two 16-byte loads per trilinear sample instead of four paired loads, the identical rfront
 * bit for bit. bricks == null: exactly nv_cast_rays / nv_cast_rays_wide. Replaces the same launch as
 * nv_cast_rays (nebula3d.cast_rays, nebula3d.py:247-270; sim.py:936-937). */
int nv_cast_rays_ex(const float* rho, const float* bricks, int nx, int ny, int nz, float sx, float sy, float sz,
                    float Q, float dr, int nsteps, float dx, float rcap, float bx, float by, float bz, float* rfront,
                    int nu, int nphi, int wide, void* stream);

/* nv_cast_rays_ex with bricks, allocating nothing: the brick array is built in a static device array of the library
 * (128 MiB per GPU, held until the process ends; grids whose nv_ray_bricks_size exceeds it are rejected with
 * NV_ERR_ARG) instead of a caller buffer, then sampled by the same casts, so rfront is bit-identical to
 * nv_ray_bricks + nv_cast_rays_ex (and to nv_cast_rays / nv_cast_rays_wide). No hipMalloc, hipFree or allocating
 * library call. The array is shared by every call on one GPU: calls that overlap there (other streams or threads)
 * race. No counterpart in the reference: it is for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc). */
int nv_cast_rays_noalloc(const float* rho, int nx, int ny, int nz, float sx, float sy, float sz, float Q, float dr,
                         int nsteps, float dx, float rcap, float bx, float by, float bz, float* rfront, int nu,
                         int nphi, int wide, void* stream);

/* Batched ray casts: every (grid, star) job in one launch, each job's rfront bit-identical to nv_cast_rays
 * (nv_cast_rays_wide when wide != 0) on that job alone. Jobs with bricks get their grid's brick array built first,
 * in one more launch over all such grids. No reference counterpart (Sim.ionize launches per star,
 * sim.py:936-937). Plan and launch as for nv_hydro_stage_batch. */
long long nv_cast_rays_batch_plan_size(int n);
int nv_cast_rays_batch_plan(const NvRayDesc* jobs, int n, void* plan, long long plan_bytes);
int nv_cast_rays_batch(const void* plan_host, const void* plan_dev, void* stream);

/* Per-cell front lookup: xi = max(xi, ramp), dfront = min(dfront, r - rf). Replaces nebula3d.apply_front
 * (nebula3d.py:281-307), launched per star by Sim.ionize (sim.py:938). */
int nv_apply_front(const float* rfront, int nu, int nphi, float sx, float sy, float sz, float dx, float width,
                   float* xi, float* dfront, int nx, int ny, int nz, void* stream);

/* dst = src/2 + (mean of the 6 clamped neighbours)/2. Replaces nebula3d.smooth_xi (nebula3d.py:310-315). */
int nv_smooth_xi(const float* src, float* dst, int nx, int ny, int nz, void* stream);

/* float32 -> float16 (round to nearest even). Replaces sim.to_half (sim.py:511-514). */
int nv_to_half(const float* src, uint16_t* dst, long long n, void* stream);

/* Fill n floats with v (xi.zero_(), df.fill_(1e3) of sim.py:930-931, smax.zero_() of sim.py:1010). */
int nv_fill_f32(float* dst, float v, long long n, void* stream);

/* The whole per-cell part of Sim.ionize (sim.py:930-943) in one pass: xi = 0 and df = 1e3, then apply_front
 * for each of the nstars stars in order (rfronts = float[nstars][nu][nphi] from nv_cast_rays), then smooth_xi
 * and the copy back into xi, then the fp16 store of dfront. Bit-identical to that launch sequence.
 * star_xyz = float[3*nstars]. dfront_h (fp16) or dfront_f (fp32, P['fp32_fields']): exactly one non-null.
 * If smax_out is non-null, the pass also reduces the CFL signal speed of (rho, m) with the new xi into
 * smax_out[0] (float bits, atomic max; the caller zeroes it); rho..mz must then be non-null. */
int nv_ionize_apply(const float* rfronts, int nu, int nphi, const float* star_xyz, int nstars, float dx,
                    float width, float* xi, uint16_t* dfront_h, float* dfront_f, int nx, int ny, int nz,
                    const float* rho, const float* mx, const float* my, const float* mz, float cn2, float ci2,
                    float* smax_out, void* stream);

/* Forms of the fused ionize pass (bits): NV_ION_VEL takes the CFL maximum over the binary16-velocity state (as
 * nv_ionize_apply_h); NV_ION_FAST evaluates each cell in fast FP32 arithmetic (the hardware square root and
 * reciprocal for the cell's distance, direction, atan2 quotient and velocities; float products for the divisions
 * by 2 pi, 6 and the ramp width): xi, dfront and the CFL maximum as close to an FP64 evaluation as the exact
 * pass's (tested on the same fronts), not bit-identical to the reference; for the fast precision forms
 * (NV_PREC_F32_FAST, NV_PREC_F16V). */
#define NV_ION_VEL 1
#define NV_ION_FAST 2

/* nv_ionize_apply with the CFL maximum taken over the binary16-velocity state (rho is not read). */
int nv_ionize_apply_h(const float* rfronts, int nu, int nphi, const float* star_xyz, int nstars, float dx,
                      float width, float* xi, uint16_t* dfront_h, float* dfront_f, int nx, int ny, int nz,
                      const float* rho, const uint16_t* vx, const uint16_t* vy, const uint16_t* vz, float cn2,
                      float ci2, float* smax_out, void* stream);

/* nv_ionize_apply / nv_ionize_apply_h of one grid given as a descriptor, in a form (NV_ION_* bits; 0 is
 * nv_ionize_apply exactly, NV_ION_VEL nv_ionize_apply_h). rf_ranges is ignored (single grids do not cull). */
int nv_ionize_apply_ex(const NvIonDesc* d, int form, void* stream);

/* Batched fused ionize pass over many grids in one launch, each member bit-identical to nv_ionize_apply_ex of
 * the same form (NV_ION_* bits: 0 is nv_ionize_apply, NV_ION_VEL nv_ionize_apply_h) on that member alone;
 * smax_out set for all members or none. No reference counterpart (sim.py:927-947 is per Sim). Plan and launch as
 * for nv_hydro_stage_batch. */
long long nv_ionize_apply_batch_plan_size(int n);
int nv_ionize_apply_batch_plan(const NvIonDesc* members, int n, int form, void* plan, long long plan_bytes);
int nv_ionize_apply_batch(const void* plan_host, const void* plan_dev, void* stream);

/* FUV dust attenuation att[nr][nu][nphi] along rays from one star. Replaces nebula3d.cast_fuv
 * (nebula3d.py:320-352), launched per star by Sim.fuv_field (sim.py:964-965). */
int nv_cast_fuv(const float* rho, const float* xi, int nx, int ny, int nz, float sx, float sy, float sz,
                float kappa, float dr, int nsteps, float dx, float binr, float bx, float by, float bz, float* att,
                int nr, int nu, int nphi, void* stream);

/* g0 += lum * trilinear(att) / max(r^2, 4e-4). Replaces nebula3d.apply_fuv (nebula3d.py:355-388). */
int nv_apply_fuv(const float* att, int nr, int nu, int nphi, float sx, float sy, float sz, float lum, float dx,
                 float binr, float* g0, int nx, int ny, int nz, void* stream);

/* The per-cell part of Sim.fuv_field (sim.py:955-978) in one pass: acc = 0, then for each star in order
 * acc += lum_s * trilinear(att_s) / max(r^2, 4e-4), then the store into g0 (fp16 g0_h or fp32 g0_f, exactly
 * one non-null). atts = float[nstars][nr][nu][nphi], star_xyz = float[3*nstars], lums = float[nstars].
 * Bit-identical to that launch sequence. */
int nv_fuv_apply(const float* atts, int nr, int nu, int nphi, const float* star_xyz, const float* lums,
                 int nstars, float dx, float binr, uint16_t* g0_h, float* g0_f, int nx, int ny, int nz,
                 void* stream);

/* ---- consumers of the step's fields ---- */

/* Timeline snapshot: the 13-byte-per-cell stored form of a state (rho, mx, my, mz, dfront, g0 binary16 saturated at
 * +-65000; xi as uint8 round(255 clamp(xi, 0, 1))). fields_half = 1: dfront and g0 are binary16 (the sim's own
 * fields), else float32. Replaces timeline.k_pack_lm / k_pack (timeline.py:19-34, 57-72). */
int nv_timeline_pack(const float* rho, const float* mx, const float* my, const float* mz, const float* xi,
                     const void* dfront, const void* g0, int fields_half, uint16_t* o_r, uint16_t* o_mx,
                     uint16_t* o_my, uint16_t* o_mz, uint8_t* o_xi, uint16_t* o_df, uint16_t* o_g0, long long n,
                     void* stream);

/* Restores a state from a snapshot (rho = max(stored rho, floor), xi = stored / 255). Replaces
 * timeline.k_unpack_lm / k_unpack (timeline.py:37-52, 75-89). */
int nv_timeline_unpack(const uint16_t* s_r, const uint16_t* s_mx, const uint16_t* s_my, const uint16_t* s_mz,
                       const uint8_t* s_xi, const uint16_t* s_df, const uint16_t* s_g0, float floor_, float* rho,
                       float* mx, float* my, float* mz, float* xi, void* dfront, void* g0, int fields_half,
                       long long n, void* stream);

/* Compression-heating proxy max(0, -div v) rho, central differences, clamped neighbours. Replaces
 * nebula3d.shock_field (nebula3d.py:182-191), behind Sim.shock. */
int nv_shock_field(const float* rho, const float* mx, const float* my, const float* mz, int nx, int ny, int nz,
                   float dx, float* out, void* stream);

/* The stored forms of one bake frame (bake.py:94-99): g0 -> min(g0, 6e4) binary16, rho -> binary16,
 * xi -> uint8 round-half-even(255 clip(xi, 0, 1)), dfront -> clip(df, -4, 4) binary16, shock -> min(sh, 6e4)
 * binary16 (shock and o_sh may both be null). fields_half as for nv_timeline_pack. */
int nv_frame_pack(const float* rho, const float* xi, const void* dfront, const void* g0, const float* shock,
                  int fields_half, uint16_t* o_rho, uint8_t* o_xi, uint16_t* o_df, uint16_t* o_g0, uint16_t* o_sh,
                  long long n, void* stream);

/* The renderer's RGBA16F volume: emission (H-alpha, [O III], [S II], dust x G0, shock) and dust per cell, faded
 * at the box faces, binary16 x 4 per cell (out = uint16[4 * cells]). Replaces render.pack_volume_lm
 * (render.py:176-204). */
int nv_pack_volume(const float* rho, const float* mx, const float* my, const float* mz, const float* xi,
                   const void* dfront, const void* g0, int fields_half, const NvPackParams* pp, int shock_on,
                   uint16_t* out, int nx, int ny, int nz, void* stream);

/* [1 4 1]/6 along one axis of a binary16 x 4 grid. Replaces render.smooth_axis (render.py:281-298). */
int nv_smooth_axis(const uint16_t* src, uint16_t* dst, int axis, int nx, int ny, int nz, void* stream);

/* ---- the pre-turbulence phase ---- */

/* Turbulent stirring: m += rho amp dt ((1 - blend) F + blend G), F and G trilinear samples of two stirring vector
 * fields (fx..fz, gx..gz: float[sx][sy][sz] each, spacing dxs) at the cell centres. Replaces nebula3d.drive
 * (nebula3d.py:165-179), launched by Sim._stir (sim.py:1030-1042). */
int nv_drive(float* rho, float* mx, float* my, float* mz, int nx, int ny, int nz, const float* fx, const float* fy,
             const float* fz, const float* gx, const float* gy, const float* gz, int sx, int sy, int sz, float blend,
             float amp, float dt, float dx, float dxs, void* stream);

/* The torus added at the end of the pre-turbulence phase: rho += torus, m *= 1 - torus / (torus + rho). Replaces
 * sim.add_torus (sim.py:353-363). */
int nv_add_torus(float* rho, float* mx, float* my, float* mz, const float* torus, long long n, void* stream);

/* out[0] += sum |m|^2 / rho, out[1] += sum rho over n cells, each cell's term in float as sim.mach_sums
 * (sim.py:502-508), the sums in FP64 (the reference's float atomics depend on their order; these do only in the
 * last bits of a double). The caller zeroes out[0..1]. */
int nv_mach_sums(const float* rho, const float* mx, const float* my, const float* mz, long long n, double* out,
                 void* stream);

/* dst = src[::2, ::2, ::2] as float32 (src binary16 when src_half). Replaces sim.every2 / every2_half
 * (sim.py:517-527), the subsampling behind the star-formation trigger. */
int nv_every2(const void* src, int src_half, float* dst, int nx, int ny, int nz, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* NEBULA_HIP_H */
