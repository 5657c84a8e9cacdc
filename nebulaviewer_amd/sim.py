"""GPU-resident simulation step with the interface of the reference's sim.Sim (sim.py:543-1222).

    sim = Sim("torus", res=64); sim.step(4); sim.update_render_fields()

Fields are torch tensors on the ROCm device, C-order [nx, ny, nz] (x, y, z; z fastest), the reference layout:
rho, mx, my, mz, xi (float32), dfront, g0 (float16, or float32 with P['fp32_fields']) and stage (4, nx, ny, nz)
float32 scratch that is free between steps (sim.py:721-733). Every kernel is a hand-written HIP kernel in
libnebula_hip.so, called through its C ABI (include/nebula_hip.h); there is no CPU fallback.

Step order (Sim._main_step, sim.py:1099-1114): ionize every ion_every steps, star weights, CFL dt, jet pulse,
two RK stages, clocks. Two fusions change no bits:
  - the per-cell part of ionize (zero/fill, apply_front per star, smooth_xi, copy, to_half) is one pass
    (nv_ionize_apply);
  - the CFL max of the next step is reduced in the epilogue of RK stage 2 (or of the ionize pass when it
    changes xi), so no separate max_speed pass reads the grid; the host still reads smax once per step,
    as the reference does (sim.py:1012), to compute dt in double.

precision: "fp32" reproduces the reference's float32 step bit for bit; "fp32-fast" uses fused multiply-adds and
hardware reciprocals / square roots in the RK stage; "fp16-vel" stores float32 density with binary16 velocity
(58 instead of 88 bytes of state traffic per cell and step) and then mx, my, mz are computed on access.

tile, ray_bricks: launch choices that change no bit: the RK stage's workgroup tile ("tune": measured per stage form
on the first stage, or an NV_TILE_* code for both) and whether the ray casts sample the density through its brick array ("auto":
measured on the first ionization, True, False); nebulaviewer_amd.autotune.

long_rays: "fp64" (default) accumulates the recombination column of a ray in FP64 when the ray's chord through
the box exceeds 512 cells, where float summation error moves the Stromgren front (nv_cast_rays_wide); shorter
rays, and every ray of grids up to 294 cells on a side, are bit-identical to the float column. "fp32" keeps the
reference's float column on every ray (nebula3d.py:262-268).
"""
import copy
import ctypes
import math
import time

import numpy as np
import torch

from . import _lib
from . import scenarios as S


def _f(x):
    import ctypes
    return ctypes.c_float(float(x))


class Sim:
    profile = False
    dt_cap = None

    PRECISIONS = ("fp32", "fp32-fast", "fp16-vel")

    IC_MODES = ("nebula3d", "matched")

    def __init__(self, scenario="torus", res=None, params=None, ic_mode="nebula3d", rays=None, fuv_grid=None,
                 ion_every=None, device="cuda:0", verbose=True, fast=True, setup=None, fused=True, seed=None,
                 precision="fp32", long_rays="fp64", tile="tune", ray_bricks="auto", device_clock=True,
                 cull_stars=True):
        # the reference's constructor (sim.py:553-565): positional order up to fast, fast=True as the viewer runs it
        # (1-voxel rays, ionization every >= 3 steps, sim.py:743, 758); bake.py and the benchmark pass fast=False
        if ic_mode not in self.IC_MODES:
            raise ValueError(f"ic_mode must be one of {self.IC_MODES}")
        if ic_mode == "matched":
            raise ValueError("ic_mode='matched' (the reference's resolution-matched noise for its artistic scenes, "
                             "sim.py:270-330) is not built: the scenarios here are the seeded nebula3d.grf3 workloads "
                             "(ic_mode='nebula3d')")
        self.ic_mode = ic_mode
        if precision not in self.PRECISIONS:
            raise ValueError(f"precision must be one of {self.PRECISIONS}")
        if long_rays not in ("fp64", "fp32"):
            raise ValueError("long_rays must be 'fp64' or 'fp32'")
        if tile != "tune" and not (isinstance(tile, int) and tile >= 0 and (tile & 0xFF) in (0, 1, 2, 3)):
            raise ValueError("tile must be 'tune' or an NV_TILE_* code (0 = the library's rule), optionally with "
                             "NV_TILE_DEPTH(n) (n << 8)")
        self.precision = precision
        self.long_rays = long_rays
        self.tile = tile
        # ray casts sample the density through its brick array (True), directly (False) or by measurement ("auto");
        # the fronts are the same bits either way
        if ray_bricks not in ("auto", True, False):
            raise ValueError("ray_bricks must be 'auto', True or False")
        self.ray_bricks = ray_bricks
        # main-phase steps whose host-side state does not follow the clock take their dt on the device (nv_cfl_dt):
        # no readback per step, the same bits (False: the host computes every dt, sim.py:1009-1012)
        self.device_clock = bool(device_clock)
        self._clk = None
        # two or more switched-on stars: the casts record each star's front range and the fused ionize pass skips,
        # per cell, the stars that provably change neither xi nor dfront (the same bits either way)
        self.cull_stars = bool(cull_stars)
        self._mplans = None
        self._mplan_key = None
        self._dirs_tab = None
        self.device = torch.device(device)
        self.fast = fast
        self.verbose = verbose
        self.rays_override = rays
        self.fuv_override = fuv_grid
        self.ion_every_override = ion_every
        self.fused = fused
        self.prof = {}
        self.timer = None         # optional timing.KernelTimer: events around each kernel group
        self._seed = seed
        _lib.load()
        if self.device.type != "cuda":
            raise _lib.NebulaHipError("Sim runs on a ROCm device (device='cuda:N')")
        self.reset(scenario, res, params, setup=setup)

    # ------------------------------------------------------------------ helpers
    def log(self, *a):
        if self.verbose:
            print("[sim]", *a, flush=True)

    def _tick(self, key, t0):
        self.prof[key] = self.prof.get(key, 0.0) + (time.perf_counter() - t0)

    @property
    def _stream(self):
        return torch.cuda.current_stream(self.device).cuda_stream

    def _t0(self, name):
        return self.timer.start(name) if self.timer is not None else None

    def _t1(self, tok):
        if tok is not None:
            self.timer.stop(tok)

    @staticmethod
    def _p(t):
        return t.data_ptr() if t is not None else None

    # ------------------------------------------------------------------ fields
    @property
    def vel_state(self):
        return self.precision == "fp16-vel"

    @property
    def ion_form(self):
        """the fused ionize pass's form (NV_ION_* bits): the fast precision forms evaluate it in fast arithmetic"""
        return (_lib.NV_ION_VEL if self.vel_state else 0) | (_lib.NV_ION_FAST if self.precision != "fp32" else 0)

    def ion_desc(self, nstars, width):
        """NvIonDesc of this grid's fused ionize pass over its first nstars rfront tables"""
        nx, ny, nz = self.shape
        half = self.dfront.dtype == torch.float16
        d = _lib.NvIonDesc()
        d.rfronts = self.rfronts.data_ptr()
        d.nu, d.nphi = self.rfronts.shape[1:]
        d.star_xyz = self.star_xyz.data_ptr()
        d.nstars = nstars
        d.dx = self.dx
        d.width = width
        d.xi = self.xi.data_ptr()
        d.dfront_h = self.dfront.data_ptr() if half else None
        d.dfront_f = None if half else self.dfront.data_ptr()
        d.nx, d.ny, d.nz = nx, ny, nz
        for c in range(4):
            d.state[c] = self.U[c].data_ptr()
        d.cn2, d.ci2 = self.cn2, self.ci2
        d.smax_out = self.smax.data_ptr()
        return d

    @property
    def rho(self):
        """density [nx, ny, nz] float32"""
        return self.U[0]

    def _mom(self, c):
        return self.U[0] * self.U[c].float() if self.vel_state else self.U[c]

    @property
    def mx(self):
        return self._mom(1)

    @property
    def my(self):
        return self._mom(2)

    @property
    def mz(self):
        return self._mom(3)

    def set_state(self, rho, mx, my, mz):
        """overwrite the conserved state from float32 tensors (timeline restores, tests)"""
        if self.vel_state:
            self.U[0].copy_(rho)
            r = self.U[0].clamp_min(self.P["floor"])
            for dst, m in zip(self.U[1:], (mx, my, mz)):
                dst.copy_(m / r)
        else:
            for dst, src in zip(self.U, (rho, mx, my, mz)):
                dst.copy_(src)
        self.fields_modified()

    # ------------------------------------------------------------------ setup
    def reset(self, scenario=None, res=None, params=None, setup=None):
        t_setup = time.perf_counter()
        for k in ("U", "xi", "dfront", "g0", "stage", "_brick_buf", "rfronts", "atts"):
            self.__dict__.pop(k, None)
        if setup is None:
            setup = S.make_setup(scenario or "torus", res, self._seed, params)
        elif params:
            setup.P.update(params)
        if setup.P.get("trigger", 0) > 0:
            # the reference forms stars where dense gas collapses (Sim._check_triggers, sim.py:1061-1097, host
            # NumPy/scipy); that decision is not built, so a run that asks for it is refused rather than run without
            raise NotImplementedError("star formation (P['trigger'] > 0) is not built")
        self.setup = setup
        self.scenario = setup.name
        self.res = setup.res
        self.P = setup.P
        P = self.P
        nx, ny, nz = self.shape = tuple(setup.shape)
        self.dx = setup.dx
        self.L = setup.L
        C = S.constants(setup, fast=self.fast, ion_every=self.ion_every_override)
        self.C = C
        self.cn2, self.ci2, self.diag, self.dr, self.nsteps = C["cn2"], C["ci2"], C["diag"], C["dr"], C["nsteps"]
        self.bmax = C["bmax"]
        self.ion_every, self.fuv_every = C["ion_every"], C["fuv_every"]
        self.stars = [tuple(float(v) for v in s) for s in setup.stars]
        self.kinds = list(setup.kinds)
        self.jets = [tuple(float(v) for v in j[:3]) for j in setup.jets]
        self.jkinds = list(setup.jkinds)
        self.cap = max(len(self.stars), P["max_stars"])

        dev = self.device
        f32 = dict(dtype=torch.float32, device=dev)
        rho = torch.from_numpy(np.ascontiguousarray(setup.rho, dtype=np.float32)).to(dev)
        z = lambda a: (torch.zeros(self.shape, **f32) if a is None
                       else torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(dev))
        m = [z(setup.mx), z(setup.my), z(setup.mz)]
        if self.vel_state:
            # float32 rho, binary16 velocity (not momentum): the state of nv_hydro_stage_h
            r = rho.clamp_min(P["floor"])
            self.U = [rho] + [(c / r).half() for c in m]
            del r, m
            # the stage block (float32 rho, binary16 velocity: 10 B/cell) is carved from the buffer that also holds
            # the density bricks of the ray casts while ionize runs (the stage is free then)
            n = math.prod(self.shape)
            self._brick_buf = torch.zeros(self._brick_floats(), **f32)
            halves = self._brick_buf.view(torch.float16)
            self.stage = [self._brick_buf[:n].view(self.shape)] + [halves[(2 + c) * n:(3 + c) * n].view(self.shape)
                                                                    for c in range(3)]
        else:
            self.U = [rho] + m
            # the float32 stage block, at the head of the (slightly larger) buffer of the ray casts' density bricks
            self._brick_buf = torch.zeros(self._brick_floats(), **f32)
            self.stage = self._brick_buf[:4 * math.prod(self.shape)].view((4,) + self.shape)
        self.xi = torch.zeros(self.shape, **f32)
        hdt = torch.float32 if P.get("fp32_fields") else torch.float16
        self.dfront = torch.full(self.shape, 1e3, dtype=hdt, device=dev)
        self.g0 = torch.zeros(self.shape, dtype=hdt, device=dev)
        self.smax = torch.zeros(1, **f32)
        self.star_arr = torch.zeros((self.cap, 4), **f32)
        self.jet_arr = torch.zeros((self.cap, 4), **f32)
        self.star_xyz = torch.zeros((self.cap, 3), **f32)
        self.lums = torch.zeros(self.cap, **f32)
        self._wkey = None
        self._xyz_key = None
        self._smax_ok = False
        self.generation = getattr(self, "generation", 0) + 1
        nu, nphi = self.rays_override or S.ray_grid(nz, self.fast)
        self.rfronts = torch.zeros((self.cap, nu, nphi), **f32)
        self.rfront = self.rfronts[0]
        fg = tuple(self.fuv_override or S.fuv_grid(nz))
        self.fuv_shape = fg
        self.atts = None          # allocated on the first FUV update (render cadence only)
        self.t = 0.0
        self.it = 0
        self.dt = 0.0
        self.version = 0
        self.rf_version = -1
        self.fuv_version = -10 ** 9
        self.events = []
        self.mach = None
        self.tp, self.ip = 0.0, 0
        self.torus = setup.torus
        self.stirring = S.Stirring(setup)
        self.sshape, self.dxs = self.stirring.shape, S.Stirring.DXS
        self.drv = None
        if P.get("pre_t", 0.0) > 0:
            # pre-turbulence phase (sim.py:786-789): stirred hydro, no stars, until tp reaches pre_t
            self.phase = "pre"
            self.drv = {"A": self._mkF(), "B": self._mkF(), "clock": 0.0}
        else:
            self.phase = "main"
            self._start_main()
        torch.cuda.synchronize(dev)
        self.setup_s = time.perf_counter() - t_setup
        self.log(f"{self.scenario} grid {nx}x{ny}x{nz} ({nx * ny * nz / 1e6:.1f}M cells) rays {nu}x{nphi} "
                 f"fuv {fg} solver {P['solver']}/{P['limiter']} setup {self.setup_s:.2f}s, {len(self.stars)} stars")

    def _start_main(self):
        """end of the turbulent phase: the torus added at rest, the first ionization (sim.py:806-818)"""
        if self.torus is not None:
            t = torch.from_numpy(np.ascontiguousarray(self.torus, dtype=np.float32)).to(self.device)
            U = self.U if not self.vel_state else None
            if U is None:
                raise ValueError("the torus of a pre-turbulence run needs the float32 state")
            _lib.call("nv_add_torus", *[u.data_ptr() for u in U], t.data_ptr(), t.numel(), self._stream)
            torch.cuda.synchronize(self.device)
            self.torus = None
            self.fields_modified()
        self.phase = "main"
        self._set_star_weights()
        self.ionize()

    # ------------------------------------------------------------------ pre-turbulence (sim.py:1030-1060)
    def _mkF(self):
        """the next stirring field (3 device arrays on the stirring grid)"""
        return [torch.from_numpy(c).to(self.device) for c in self.stirring.next()]

    def _stir(self, dt, amp):
        d = self.drv
        blend = min(d["clock"] / self.P["pre_corr"], 1.0)
        U = self.U
        if self.vel_state:
            raise ValueError("stirring needs the float32 state")
        _lib.call("nv_drive", *[u.data_ptr() for u in U], *self.shape, *[f.data_ptr() for f in d["A"]],
                  *[f.data_ptr() for f in d["B"]], *self.sshape, _f(blend), _f(amp), _f(dt), _f(self.dx),
                  _f(self.dxs), self._stream)
        d["clock"] += dt
        if d["clock"] >= self.P["pre_corr"]:
            d["A"], d["B"], d["clock"] = d["B"], self._mkF(), 0.0
        self._smax_ok = False          # the stage's fused CFL max predates the stirring

    def mach_number(self):
        """mass-weighted rms Mach number sqrt(sum m^2/rho / sum rho) / c_n (sim.py:1052-1055), sums in FP64"""
        s = torch.zeros(2, dtype=torch.float64, device=self.device)
        _lib.call("nv_mach_sums", *[t.data_ptr() for t in (self.rho, self.mx, self.my, self.mz)], self.rho.numel(),
                  s.data_ptr(), self._stream)
        a, b = s.tolist()
        return math.sqrt(a / b) * self.P["ci_over_cn"]

    def _pre_step(self):
        dt = self._compute_dt()
        self._hydro(dt, 0, 0.0, 0.0)
        self._stir(dt, self.P["pre_amp"])
        self.tp += dt
        self.ip += 1
        self.dt = dt
        if self.tp >= self.P["pre_t"]:
            self.mach = self.mach_number()
            self.log(f"pre-turbulence done: {self.ip} steps t={self.tp:.3f} mass-weighted Mach ~{self.mach:.1f}")
            if self.P["drive_main"] <= 0:
                self.drv = None
            self._start_main()

    # ------------------------------------------------------------------ stars (sim.py:879-925)
    def q_of(self, i):
        s = self.stars[i]
        return s[3] * self.P["q_main"] if self.kinds[i] == "main" else s[3]

    def ramp(self, i):
        age = self.t - self.stars[i][4]
        if self.phase != "main" or age < 0.0:
            return 0.0
        tr = float(self.P.get("q_ramp", 0.0))
        if tr <= 0.0 or age >= tr:
            return 1.0
        x = age / tr
        s = x * x * (3.0 - 2.0 * x)
        return max(s * s * s, 1e-4)

    def q_eff(self, i):
        return self.q_of(i) * self.ramp(i)

    def jet_of(self, i):
        k = self.jkinds[i]
        return self.P["jet"] if k == "jet" else (self.P["jet_proto"] if k == "proto" else 0.0)

    def active_stars(self):
        out = []
        for i, s in enumerate(self.stars):
            if self.t >= s[4]:
                out.append((s[0], s[1], s[2], self.q_eff(i) / self.P["q_main"], self.t - s[4]))
        return out

    def _set_star_weights(self):
        qm = self.P["q_main"]
        w = [(s[0], s[1], s[2], (self.q_eff(i) / qm) if self.t >= s[4] else 0.0) for i, s in enumerate(self.stars)]
        w += [(0.0, 0.0, 0.0, 0.0)] * (self.cap - len(w))
        key = tuple(round(v[3], 9) for v in w) + tuple(self.jet_of(i) for i in range(len(self.stars)))
        if key == self._wkey:
            return
        self._wkey = key
        self.star_arr.copy_(torch.tensor(np.array(w, dtype=np.float32)))
        jw = [(a[0], a[1], a[2], self.jet_of(i)) for i, a in enumerate(self.jets)]
        jw += [(0.0, 0.0, 1.0, 0.0)] * (self.cap - len(jw))
        self.jet_arr.copy_(torch.tensor(np.array(jw, dtype=np.float32)))

    def _culled_plans(self, on, use_bricks, width):
        """(ray plan, ionize plan) of the batch entry points for this grid's switched-on stars: one ray job per star
        (its front table, the shared direction table, its front range) and the fused ionize pass over the grid with
        the stars' front ranges, so that a cell skips the stars that provably change neither xi nor dfront. The
        casts are cast_one on the same rays and the pass the same per-cell arithmetic, so xi, dfront and the CFL
        maximum are those of the per-star casts and nv_ionize_apply_ex bit for bit. Rebuilt when a star, its
        luminosity, the brick choice or the grid's fields change."""
        from .batch import _Plan
        nx, ny, nz = self.shape
        nu, nphi = self.rfronts.shape[1:]
        key = (tuple(on), tuple(self.q_eff(i) for i in on), tuple(tuple(self.stars[i][:3]) for i in on), use_bricks,
               self.ion_form, self.generation, self.long_rays, width, self.rfronts.data_ptr(), self.star_xyz.data_ptr())
        if key == self._mplan_key:
            return self._mplans
        L = _lib.load()
        if self._dirs_tab is None or self._dirs_tab.numel() != 4 * nu * nphi:
            self._dirs_tab = torch.empty(4 * nu * nphi, dtype=torch.float32, device=self.device)
            _lib.call("nv_ray_dirs", nu, nphi, self._dirs_tab.data_ptr(), self._stream)
        per = int(L.nv_ray_range_floats(nu, nphi))
        ranges = torch.zeros(per * len(on), dtype=torch.float32, device=self.device)
        Lx, Ly, Lz = self.bmax
        jobs = (_lib.NvRayDesc * len(on))()
        for slot, i in enumerate(on):
            j = jobs[slot]
            j.rho = self.U[0].data_ptr()
            j.bricks = self._brick_buf.data_ptr() if use_bricks else None
            j.nx, j.ny, j.nz = nx, ny, nz
            j.sx, j.sy, j.sz = self.stars[i][:3]
            j.Q = self.q_eff(i)
            j.dr, j.nsteps, j.dx, j.rcap = self.dr, self.nsteps, self.dx, self.diag
            j.bx, j.by, j.bz = Lx, Ly, Lz
            j.rfront = self.rfronts[slot].data_ptr()
            j.nu, j.nphi = nu, nphi
            j.wide = 1 if self.long_rays == "fp64" else 0
            j.dirs = self._dirs_tab.data_ptr()
            j.rf_range = ranges.data_ptr() + 4 * per * slot
        ion = (_lib.NvIonDesc * 1)()
        ion[0] = self.ion_desc(len(on), width)
        ion[0].rf_ranges = ranges.data_ptr()
        P = _lib.P
        form = self.ion_form

        def build_rays(buf, nbytes):
            _lib.call("nv_cast_rays_batch_plan", ctypes.cast(jobs, P), len(on), buf, nbytes)

        def build_ion(buf, nbytes):
            _lib.call("nv_ionize_apply_batch_plan", ctypes.cast(ion, P), 1, form, buf, nbytes)
        self._mplans = (_Plan("nv_cast_rays_batch_plan_size", build_rays, len(on), self.device),
                        _Plan("nv_ionize_apply_batch_plan_size", build_ion, 1, self.device), ranges)
        self._mplan_key = key
        return self._mplans

    def _upload_xyz(self, on):
        """star_xyz[:len(on)] = the positions of the switched-on stars; the copy (from pageable host memory, which
        waits for the stream) only when they changed"""
        xyz = np.array([self.stars[i][:3] for i in on], dtype=np.float32).reshape(-1, 3)
        key = xyz.tobytes()
        if len(on) and key != self._xyz_key:
            self.star_xyz[:len(on)].copy_(torch.from_numpy(xyz))
        self._xyz_key = key

    def replace_stars(self, stars, kinds, jets, jkinds):
        self.stars = [tuple(float(v) for v in s) for s in stars]
        self.kinds = list(kinds)
        self.jets = [tuple(float(v) for v in j[:3]) for j in jets]
        self.jkinds = list(jkinds)
        if len(self.stars) > self.cap:
            self.cap = len(self.stars)
            f32 = dict(dtype=torch.float32, device=self.device)
            self.star_arr = torch.zeros((self.cap, 4), **f32)
            self.jet_arr = torch.zeros((self.cap, 4), **f32)
            self.star_xyz = torch.zeros((self.cap, 3), **f32)
            self.lums = torch.zeros(self.cap, **f32)
            self.rfronts = torch.zeros((self.cap,) + tuple(self.rfronts.shape[1:]), **f32)
            self.atts = None
        self._wkey = None
        self._xyz_key = None
        self._set_star_weights()
        self.ionize()

    def stars_absolute(self):
        st = [(s[0], s[1], s[2], self.q_of(i), s[4]) for i, s in enumerate(self.stars)]
        jt = [(j[0], j[1], j[2], self.jet_of(i)) for i, j in enumerate(self.jets)]
        return st, jt

    # ------------------------------------------------------------------ physics pieces
    def ionize(self):
        """Stromgren fronts of every switched-on star -> xi, dfront (Sim.ionize, sim.py:927-947)."""
        t0 = time.perf_counter()
        nx, ny, nz = self.shape
        nu, nphi = self.rfronts.shape[1:]
        strm = self._stream
        on = [i for i, s in enumerate(self.stars) if self.t >= s[4]]
        Lx, Ly, Lz = self.bmax
        # The casts sample the density either directly or from its brick array (two 16-byte loads per sample),
        # built in the 16 B/cell buffer of the RK stage block, which is free while ionize runs; the same fronts bit
        # for bit. ray_bricks="auto" measures both on the first ionization (nebulaviewer_amd.autotune).
        def casts(use_bricks, timed=False):
            bricks = None
            if use_bricks:
                bricks = self._p(self._brick_buf)
                tk = self._t0("ray_bricks") if timed else None
                _lib.call("nv_ray_bricks", self._p(self.U[0]), nx, ny, nz, bricks, strm)
                self._t1(tk)
            tk = self._t0("cast_rays") if timed else None
            for slot, i in enumerate(on):
                sx, sy, sz = self.stars[i][:3]
                _lib.call("nv_cast_rays_ex", self._p(self.U[0]), bricks, nx, ny, nz,
                          _f(sx), _f(sy), _f(sz), _f(self.q_eff(i)),
                          _f(self.dr), self.nsteps, _f(self.dx), _f(self.diag), _f(Lx), _f(Ly), _f(Lz),
                          self._p(self.rfronts[slot]), nu, nphi, 1 if self.long_rays == "fp64" else 0, strm)
            self._t1(tk)

        use = self.ray_bricks
        if not on:
            use = False
        elif use == "auto":
            from . import autotune
            use = autotune.tune_rays(self, len(on), casts)
        self.ray_bricks_used = bool(use)
        width = 1.2 * self.dx
        half = self.dfront.dtype == torch.float16
        plans = self._culled_plans(on, bool(use), width) if self.fused and self.cull_stars and len(on) >= 2 else None
        if plans is None:
            casts(bool(use), timed=True)
        if self.fused:
            self._upload_xyz(on)
            self.smax.zero_()
            if plans is not None:
                # the batch entry points with this one grid: every star's casts in one launch (the brick build
                # inside it), each recording its front range, then the culled pass
                tk = self._t0("cast_rays")
                _lib.call("nv_cast_rays_batch", *plans[0].args, strm)
                self._t1(tk)
                tk = self._t0("ionize_apply")
                _lib.call("nv_ionize_apply_batch", *plans[1].args, strm)
                self._t1(tk)
            else:
                d = self.ion_desc(len(on), width)
                tk = self._t0("ionize_apply")
                _lib.call("nv_ionize_apply_ex", ctypes.byref(d), self.ion_form, strm)
                self._t1(tk)
            self._smax_ok = True
        else:
            # the reference launch sequence, kernel by kernel (sim.py:930-943)
            if self.vel_state:
                raise ValueError("the unfused ionize sequence needs the float32 state (precision='fp32')")
            n = nx * ny * nz
            df = self.stage[1]
            _lib.call("nv_fill_f32", self._p(self.xi), _f(0.0), n, strm)
            _lib.call("nv_fill_f32", self._p(df), _f(1e3), n, strm)
            for slot, i in enumerate(on):
                sx, sy, sz = self.stars[i][:3]
                _lib.call("nv_apply_front", self._p(self.rfronts[slot]), nu, nphi, _f(sx), _f(sy), _f(sz),
                          _f(self.dx), _f(width), self._p(self.xi), self._p(df), nx, ny, nz, strm)
            _lib.call("nv_smooth_xi", self._p(self.xi), self._p(self.stage[0]), nx, ny, nz, strm)
            self.xi.copy_(self.stage[0])
            if half:
                _lib.call("nv_to_half", self._p(df), self._p(self.dfront), n, strm)
            else:
                self.dfront.copy_(df)
            self._smax_ok = False
        if self.profile:
            torch.cuda.synchronize(self.device)
            self._tick("ionize", t0)

    def fuv_field(self):
        """G0 = sum over stars of L exp(-tau_FUV) / r^2 (Sim.fuv_field, sim.py:949-978), stored into g0."""
        nx, ny, nz = self.shape
        fnr, fnu, fnp = self.fuv_shape
        if self.atts is None:
            self.atts = torch.zeros((self.cap, fnr, fnu, fnp), dtype=torch.float32, device=self.device)
        self.att = self.atts[0]
        binr = self.diag / fnr
        strm = self._stream
        on = [i for i, s in enumerate(self.stars) if self.t >= s[4]]
        Lx, Ly, Lz = self.bmax
        for slot, i in enumerate(on):
            sx, sy, sz = self.stars[i][:3]
            _lib.call("nv_cast_fuv", self._p(self.U[0]), self._p(self.xi),
                      nx, ny, nz, _f(sx), _f(sy), _f(sz),
                      _f(self.P["kappa_fuv"]), _f(self.dr), self.nsteps, _f(self.dx), _f(binr), _f(Lx), _f(Ly),
                      _f(Lz), self._p(self.atts[slot]), fnr, fnu, fnp, strm)
        lum = np.array([self.q_eff(i) / self.P["q_main"] for i in on], dtype=np.float32)
        self._upload_xyz(on)
        if len(on):
            self.lums[:len(on)].copy_(torch.from_numpy(lum))
        half = self.g0.dtype == torch.float16
        _lib.call("nv_fuv_apply", self._p(self.atts), fnr, fnu, fnp, self._p(self.star_xyz), self._p(self.lums),
                  len(on), _f(self.dx), _f(binr), self._p(self.g0) if half else None,
                  None if half else self._p(self.g0), nx, ny, nz, strm)

    def update_render_fields(self, force=False):
        """recompute G0 at the render cadence (Sim.update_render_fields, sim.py:989-1007)"""
        if self.rf_version == self.version and not force:
            return False
        n_on = sum(1 for s in self.stars if self.t >= s[4])
        tr = float(self.P.get("q_ramp", 0.0))
        ramping = tr > 0 and any(0.0 <= self.t - s[4] < tr for s in self.stars)
        if ramping and self.version - self.fuv_version >= max(2, self.fuv_every // 4):
            force = True
        if (force or self.version - self.fuv_version >= self.fuv_every
                or n_on != getattr(self, "fuv_nstars", -1) or self.phase != getattr(self, "fuv_phase", None)):
            self.fuv_field()
            self.fuv_version, self.fuv_nstars, self.fuv_phase = self.version, n_on, self.phase
        self.rf_version = self.version
        return True

    def fields_modified(self):
        """Call after writing rho/m/xi from outside (timeline restore, tests): the cached CFL max is stale."""
        self._smax_ok = False
        self.version += 1

    def _max_speed(self):
        """the CFL signal speed of the current state into smax (nebula3d.max_speed)"""
        nx, ny, nz = self.shape
        self.smax.zero_()
        if self.vel_state:
            _lib.call("nv_max_speed_h", *[self._p(t) for t in self.U[1:]], self._p(self.xi), nx, ny, nz,
                      _f(self.cn2), _f(self.ci2), self._p(self.smax), self._stream)
        else:
            _lib.call("nv_max_speed", *[self._p(t) for t in self.U], self._p(self.xi), nx, ny, nz, _f(self.cn2),
                      _f(self.ci2), self._p(self.smax), self._stream)

    def _compute_dt(self):
        """dt = cfl dx / max(|u|+|v|+|w|+c) (Sim._compute_dt, sim.py:1009-1012), the max read in double."""
        if not self._smax_ok:
            self._max_speed()
        smax = float(self.smax.item())
        self._smax_ok = False
        return self.P["cfl"] * self.dx / smax

    def _brick_floats(self):
        """floats of the buffer that holds the RK stage block and, while ionize runs, the density bricks (16 B per
        cell plus one ghost layer: nv_ray_bricks_size)"""
        nx, ny, nz = self.shape
        return max(4 * nx * ny * nz, int(_lib.load().nv_ray_bricks_size(nx, ny, nz)) // 4)

    def _solver_lim(self):
        solver = S.SOLVERS[self.P.get("solver", "rusanov")]
        return solver, (S.LIMITERS[self.P.get("limiter", "minmod")] if solver == 1 else 0)

    def _stage_desc(self, src, u0, dst, a, b, dt, nstars, wind, pulse, smax_ptr, dt_dev=None):
        """the arguments of one RK stage (nv_hydro_stage, sim.py:1017-1028) as an NvStageDesc; float fields are
        rounded to float32 as a Warp launch rounds its scalar arguments"""
        P = self.P
        d = _lib.NvStageDesc()
        for c in range(4):
            d.u[c], d.u0[c], d.out[c] = src[c], u0[c], dst[c]
        d.xi = self._p(self.xi)
        d.stars, d.jets = self._p(self.star_arr), self._p(self.jet_arr)
        d.smax_out = smax_ptr
        d.dt_dev = dt_dev
        d.nx, d.ny, d.nz = self.shape
        d.nstars = nstars
        d.wind, d.jet_cos, d.jet_r, d.pulse = wind, self.C["jet_cos"], P["jet_r"], pulse
        d.dt, d.dx, d.a, d.b = dt, self.dx, a, b
        d.floor_, d.cn2, d.ci2, d.vmax = P["floor"], self.cn2, self.ci2, 6.0
        return d

    def _hydro(self, dt, nstars, wind, pulse, dt_dev=None):
        """Two RK stages (Sim._hydro, sim.py:1014-1028): U1 = U + dt L(U); U = U/2 + (U1 + dt L(U1))/2 in place.
        Stage 2 also reduces the next step's CFL signal speed into self.smax. The workgroup tile is the measured
        fastest for this grid (tile="tune", nebulaviewer_amd.autotune) or the one given; no tile changes a bit.
        dt_dev: a device float the stages read their dt from (the device clock), dt then unused."""
        import ctypes
        solver, lim = self._solver_lim()
        if self.tile == "tune":
            from . import autotune
            self.tile_shape = autotune.tune_sim(self)          # one tile per stage form
        else:
            self.tile_shape = (self.tile, self.tile)
        code = {"fp32": 0, "fp32-fast": 1, "fp16-vel": 2}[self.precision]
        strm = self._stream
        U = [self._p(t) for t in self.U]
        U1 = [self._p(self.stage[c]) for c in range(4)]
        d1 = self._stage_desc(U, U, U1, 0.0, 1.0, dt or 0.0, nstars, wind, pulse, None, dt_dev)
        d2 = self._stage_desc(U1, U, U, 0.5, 0.5, dt or 0.0, nstars, wind, pulse, self._p(self.smax), dt_dev)
        tk = self._t0("hydro_stage")
        _lib.call("nv_hydro_stage_ex", ctypes.byref(d1), code, solver, lim, self.tile_shape[0], strm)
        self._t1(tk)
        self.smax.zero_()
        tk = self._t0("hydro_stage")
        _lib.call("nv_hydro_stage_ex", ctypes.byref(d2), code, solver, lim, self.tile_shape[1], strm)
        self._t1(tk)
        self._smax_ok = True

    def _main_step(self):
        P = self.P
        if self.it % self.ion_every == 0:
            self.ionize()
        self._set_star_weights()
        dt = self._compute_dt()
        if self.dt_cap is not None:
            dt = min(dt, self.dt_cap)
        pulse = max(0.0, 1.0 + P["jet_pulse"] * math.sin(2 * math.pi * self.t / P["jet_period"]))
        self._hydro(dt, len(self.stars), P["wind"], pulse)
        if P["drive_main"] > 0 and self.drv is not None:
            self._stir(dt, P["pre_amp"] * P["drive_main"])
        self.t += dt
        self.it += 1
        self.dt = dt

    def _clock_follows_host(self):
        """True when host-side state of the main-phase step changes with the clock over the coming steps: a star
        still to switch on, a luminosity ramp in progress, a pulsed active jet, main-phase stirring, or a dt cap
        the device clock does not take (<= 0)"""
        tr = float(self.P.get("q_ramp", 0.0))
        for i, s in enumerate(self.stars):
            if self.t < s[4] or (tr > 0.0 and self.t - s[4] < tr):
                return True
            if self.jet_of(i) > 0.0 and self.P["jet_pulse"] != 0.0:
                return True
        if self.P["drive_main"] > 0 and self.drv is not None:
            return True
        return self.dt_cap is not None and not self.dt_cap > 0.0

    def _device_steps(self, k):
        """k main-phase steps (Sim._main_step, sim.py:1099-1114) with the clock on the device: nv_cfl_dt takes dt
        from the CFL maximum the previous stage 2 (or the ionize pass) reduced, in the host's double arithmetic
        (cfl dx / smax, the cap, the float the stages read, t += dt), so no step waits for a readback; t and dt
        return to the host after the last step. Only for steps whose host-side state does not follow the clock
        (_clock_follows_host), so the star tables and the jet pulse are those of the first step."""
        P = self.P
        c = self._clk
        if c is None:
            f64 = dict(dtype=torch.float64, device=self.device)
            c = self._clk = {"cfl_dx": torch.zeros(1, **f64), "dt_cap": torch.zeros(1, **f64),
                             "dt_d": torch.zeros(1, **f64), "t": torch.zeros(1, **f64),
                             "dt_f": torch.zeros(1, dtype=torch.float32, device=self.device)}
        c["cfl_dx"].fill_(P["cfl"] * self.dx)
        c["dt_cap"].fill_(self.dt_cap if self.dt_cap is not None else 0.0)
        c["t"].fill_(self.t)
        self._set_star_weights()
        pulse = max(0.0, 1.0 + P["jet_pulse"] * math.sin(2 * math.pi * self.t / P["jet_period"]))
        strm = self._stream
        for _ in range(k):
            if self.it % self.ion_every == 0:
                self.ionize()
            if not self._smax_ok:
                self._max_speed()
            _lib.call("nv_cfl_dt", self._p(self.smax), c["cfl_dx"].data_ptr(), c["dt_cap"].data_ptr(),
                      c["dt_f"].data_ptr(), c["dt_d"].data_ptr(), c["t"].data_ptr(), 1, strm)
            self._smax_ok = False
            self._hydro(None, len(self.stars), P["wind"], pulse, dt_dev=c["dt_f"].data_ptr())
            self.it += 1
            self.version += 1
        self.prof["device_clock_steps"] = self.prof.get("device_clock_steps", 0) + k
        self.t = float(c["t"].item())
        self.dt = float(c["dt_d"].item())

    # ------------------------------------------------------------------ public API
    def step(self, k=1):
        """advance k steps. Returns wall seconds."""
        t0 = time.perf_counter()
        done = 0
        while done < k:
            if self.phase == "main" and self.device_clock and not self._clock_follows_host():
                self._device_steps(k - done)
                break
            if self.phase == "pre":
                self._pre_step()
            else:
                self._main_step()
            self.version += 1
            done += 1
        self._tick("step", t0)
        self.prof["n_steps"] = self.prof.get("n_steps", 0) + k
        return time.perf_counter() - t0

    def set_param(self, name, value):
        if name == "trigger" and value > 0:
            raise NotImplementedError("star formation (P['trigger'] > 0) is not built")
        old = self.P.get(name)
        if isinstance(old, int) and not isinstance(old, bool):
            value = int(round(value))
        self.P[name] = value
        if name == "drive_main" and value > 0 and self.drv is None and self.phase == "main":
            self.drv = {"A": self._mkF(), "B": self._mkF(), "clock": 0.0}
        if name in S.LIVE:
            self.events.append((float(self.t if self.phase == "main" else 0.0), name, value))
            if name in ("q_main", "jet", "jet_proto"):
                self._wkey = None
        return value

    def tau(self):
        """timeline time: tp - pre_t (< 0) during the pre-turbulence phase, t afterwards (sim.py:1169-1171)"""
        return self.t if self.phase == "main" else self.tp - self.P["pre_t"]

    def time_label(self):
        """the viewer's clock label (sim.py:1213-1216)"""
        if self.phase == "pre":
            return f"pre-turb {self.tp:.3f}/{self.P['pre_t']:.2f}"
        return f"t={self.t:.4f}"

    def get_state_meta(self):
        """everything but the big fields needed to resume from this moment (sim.py:1173-1184): clocks, phase, stars,
        parameters, the stirring fields in use (device arrays, replaced - never modified - by the stirring) and
        the stirring stream's state"""
        d = self.drv
        return dict(t=self.t, it=self.it, tp=self.tp, ip=self.ip, dt=self.dt, phase=self.phase, mach=self.mach,
                    stars=list(self.stars), kinds=list(self.kinds), jets=list(self.jets), jkinds=list(self.jkinds),
                    P=dict(self.P), n_events=len(self.events),
                    drv=None if d is None else {"A": list(d["A"]), "B": list(d["B"]), "clock": d["clock"]},
                    rng=copy.deepcopy(self.stirring.rng.bit_generator.state),
                    torus=self.torus)

    def set_state_meta(self, m):
        """after the fields were restored (Timeline.restore; sim.py:1186-1211)"""
        for k in ("t", "it", "dt", "phase"):
            setattr(self, k, m[k])
        self.tp, self.ip, self.mach = m.get("tp", 0.0), m.get("ip", 0), m.get("mach")
        self.stars, self.kinds = list(m["stars"]), list(m["kinds"])
        self.jets, self.jkinds = list(m["jets"]), list(m["jkinds"])
        self.P = dict(m["P"])
        self.events = self.events[:m["n_events"]]
        d = m.get("drv")
        self.drv = None if d is None else {"A": list(d["A"]), "B": list(d["B"]), "clock": d["clock"]}
        if "rng" in m:
            self.stirring.rng.bit_generator.state = copy.deepcopy(m["rng"])
        self.torus = m.get("torus")
        self._wkey = None
        if self.phase == "main":
            self._set_star_weights()
        self.fields_modified()
        # xi, dfront and g0 come with the restored fields: no ray casting needed to show this state
        self.rf_version = self.version
        self.fuv_version = self.version
        self.fuv_nstars = sum(1 for s in self.stars if self.t >= s[4])
        self.fuv_phase = self.phase

    @property
    def shock(self):
        """compression heating proxy max(0, -div v) rho in a new float32 grid (nv_shock_field; sim.py:827-833)"""
        out = torch.empty(self.shape, dtype=torch.float32, device=self.device)
        rho, mx, my, mz = self.rho, self.mx, self.my, self.mz
        _lib.call("nv_shock_field", rho.data_ptr(), mx.data_ptr(), my.data_ptr(), mz.data_ptr(), *self.shape,
                  _f(self.dx), out.data_ptr(), self._stream)
        return out

    def render_scratch(self):
        """two (nx, ny, nz, 4) float16 views of the stage block, free between steps (sim.py:821-825)"""
        if self.vel_state:
            raise NotImplementedError("render_scratch needs the float32 stage block (precision='fp32')")
        flat = self.stage.view(-1).view(torch.float16)
        n = self.shape[0] * self.shape[1] * self.shape[2]
        return flat[:4 * n].view(self.shape + (4,)), flat[4 * n:8 * n].view(self.shape + (4,))

    def stats(self):
        r = self.U[0]
        return dict(ion=float(self.xi.mean()), rho_max=float(r.max()), rho_min=float(r.min()))
