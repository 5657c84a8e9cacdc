"""Many independent simulations advanced together: one launch per RK stage, per ionization pass and per CFL
step for the whole batch.

    batch = SimBatch(scenarios.batch_setups(64))
    batch.step(200)
    batch.members[3].rho          # each member is a Sim with its own fields

The reference steps one grid per Sim (sim.py:1099-1114); there is no batching in it. Member m of a SimBatch
reproduces Sim(setups[m]) stepped alone bit for bit: the batched kernels (nv_hydro_stage_batch,
nv_cast_rays_batch, nv_ionize_apply_batch) run the single-grid arithmetic on each member's tiles, and the
timestep is computed from each member's own CFL maximum in the host's double arithmetic.

Two ways to take a step:
  - device clock (when no member's host-side state depends on time during the steps: every star already on, no
    luminosity ramp in progress, no pulsed jet): nv_cfl_dt computes each member's dt and clock on the GPU, so a
    step is a fixed sequence of launches with no host round trip. With graph=True that sequence is captured
    once per kind of step (with and without ionization) as a HIP graph and replayed.
  - host clock (otherwise): one readback of the batch's CFL maxima per step, the dt computed per member on the
    host as Sim._compute_dt does, and the star tables refreshed as Sim._main_step does.
"""
import ctypes
import math

import numpy as np
import torch

from . import _lib
from . import scenarios as S
from .sim import Sim, _f

PRECISION_CODES = {"fp32": 0, "fp32-fast": 1, "fp16-vel": 2}


class _Plan:
    """A plan blob built on the host and its device copy (kept alive together)."""

    def __init__(self, size_fn, build, n, device):
        nbytes = int(getattr(_lib.load(), size_fn)(n))
        self.host = ctypes.create_string_buffer(nbytes)
        build(self.host, nbytes)
        self.dev = torch.frombuffer(bytearray(self.host.raw), dtype=torch.uint8).to(device)

    @property
    def args(self):
        return ctypes.addressof(self.host), self.dev.data_ptr()


class SimBatch:
    def __init__(self, setups, precision="fp32", device="cuda:0", rays=None, ion_every=None, graph=False,
                 verbose=False, tile="tune", dispatch="tune", queue_wgs=0, cull_stars=True):
        if not setups:
            raise ValueError("a batch needs at least one member")
        # members with two or more stars: the casts record each star's front range and the ionize pass skips, per
        # cell, the stars that provably change neither xi nor dfront (the same bits either way)
        self.cull_stars = bool(cull_stars)
        self.device = torch.device(device)
        self.precision = precision
        # workgroup tile shape of each stage form: "tune" (measured per stage, nebulaviewer_amd.autotune), 0 (the
        # library's rule), or 1 / 2 / 3 (64x8, 32x16, 64x16) for both; no shape changes a bit of the result
        self.tune = tile == "tune"
        self.tile = {1: 0, 2: 0} if self.tune else {1: tile, 2: tile}
        self.ray_bricks = "auto"  # ray casts through density bricks: True, False or measured ("auto")
        # how the stage's tiles reach the CUs, per stage form: "hardware" (one workgroup per tile), "queue" (a
        # persistent grid fed by a tile work queue, HLL + minmod only) or measured ("tune"); queue_wgs: resident
        # workgroups per CU of the persistent grid (its waves per CU over the tile's wavefronts; 0: as many as the
        # kernel's registers and LDS allow), measured with the dispatch form when that is tuned
        if dispatch not in ("tune", "hardware", "queue"):
            raise ValueError("dispatch must be 'tune', 'hardware' or 'queue'")
        if int(queue_wgs) < 0:
            raise ValueError("queue_wgs must be >= 0")
        self.dispatch = dispatch if dispatch == "tune" else {1: dispatch == "queue", 2: dispatch == "queue"}
        self.queue_wgs = {1: int(queue_wgs), 2: int(queue_wgs)}
        self.members = [Sim(fast=False, setup=st, verbose=verbose, precision=precision, device=device, rays=rays,
                            ion_every=ion_every) for st in setups]
        m0 = self.members[0]
        self._check_members()
        for m in self.members:
            if (m.P["solver"], m.P["limiter"]) != (m0.P["solver"], m0.P["limiter"]):
                raise ValueError("batch members must share the solver and limiter")
            if m.ion_every != m0.ion_every:
                raise ValueError("batch members must share the ionization cadence")
        self.n = len(self.members)
        self.ion_every = m0.ion_every
        self.solver = S.SOLVERS[m0.P.get("solver", "rusanov")]
        self.lim = S.LIMITERS[m0.P.get("limiter", "minmod")] if self.solver == 1 else 0
        dev = self.device
        self.smax = torch.zeros(self.n, dtype=torch.float32, device=dev)
        self.dt_f = torch.zeros(self.n, dtype=torch.float32, device=dev)
        self.dt_d = torch.zeros(self.n, dtype=torch.float64, device=dev)
        self.t_dev = torch.tensor([m.t for m in self.members], dtype=torch.float64, device=dev)
        self.cfl_dx = torch.tensor([m.P["cfl"] * m.dx for m in self.members], dtype=torch.float64, device=dev)
        self.dt_cap = torch.tensor([m.dt_cap or 0.0 for m in self.members], dtype=torch.float64, device=dev)
        self.it = m0.it
        self.graph = graph
        self._graphs = {}
        self._ion_key = None
        self._dirs = {}            # (nu, nphi) -> direction table of that ray grid (nv_ray_dirs)
        self._ranges = {}          # member -> its stars' front ranges (2 floats per star slot)
        self._stage_key = None
        self.timer = None
        # the CFL maximum of the current state, as Sim._compute_dt takes it when no fused value is cached
        for s, m in enumerate(self.members):
            slot = self.smax.data_ptr() + 4 * s
            nx, ny, nz = m.shape
            if m.vel_state:
                _lib.call("nv_max_speed_h", *[t.data_ptr() for t in m.U[1:]], m.xi.data_ptr(), nx, ny, nz, _f(m.cn2),
                          _f(m.ci2), slot, self._stream)
            else:
                _lib.call("nv_max_speed", *[t.data_ptr() for t in m.U], m.xi.data_ptr(), nx, ny, nz, _f(m.cn2),
                          _f(m.ci2), slot, self._stream)
        torch.cuda.synchronize(dev)

    # ------------------------------------------------------------------ helpers
    def _check_members(self):
        """The batch runs the main-phase step only (Sim._main_step, sim.py:1099-1114): a member still in its
        pre-turbulence phase (Sim._pre_step) or stirred in the main phase (drive_main > 0, sim.py:1109) would
        silently diverge from Sim, so such members are refused."""
        for s, m in enumerate(self.members):
            if m.phase != "main":
                raise ValueError(f"batch member {s} is in its '{m.phase}' phase: SimBatch steps main-phase members")
            if m.P.get("drive_main", 0.0) > 0:
                raise ValueError(f"batch member {s} stirs in the main phase (drive_main > 0): not batched")

    @property
    def _stream(self):
        return torch.cuda.current_stream(self.device).cuda_stream

    def _t0(self, name):
        return self.timer.start(name) if self.timer is not None else None

    def _t1(self, tok):
        if tok is not None:
            self.timer.stop(tok)

    @property
    def cells(self):
        return sum(m.shape[0] * m.shape[1] * m.shape[2] for m in self.members)

    def _on(self, m):
        return [i for i, s in enumerate(m.stars) if m.t >= s[4]]

    def device_clock_ok(self):
        """True when no member's host-side state changes with time over the next steps."""
        for m in self.members:
            tr = float(m.P.get("q_ramp", 0.0))
            for i, s in enumerate(m.stars):
                if m.t < s[4]:
                    return False                      # a star still to switch on
                if tr > 0.0 and m.t - s[4] < tr:
                    return False                      # luminosity ramp in progress
                if m.jet_of(i) > 0.0 and m.P["jet_pulse"] != 0.0:
                    return False                      # pulsed jet: the pulse follows the clock
        return True

    @staticmethod
    def _pulse(m):
        """the jet pulse of Sim._main_step (sim.py:1108); 1 for a member without an active jet, whose stage never
        reads it, so that its plans (and captured graphs) do not change with the clock"""
        if not any(m.jet_of(i) > 0.0 for i in range(len(m.stars))):
            return 1.0
        return max(0.0, 1.0 + m.P["jet_pulse"] * math.sin(2 * math.pi * m.t / m.P["jet_period"]))

    # ------------------------------------------------------------------ plans
    def _stage_plan(self, stage, pulses, scratch=None, smax=None):
        """the plan of RK stage 1 (U -> U1) or stage 2 (U1, U0 = U -> U, with the fused CFL maxima into smax);
        scratch: per member, four output pointers that take the result instead (the dispatch tuner's timing runs:
        stage 2 then reads U with U0 = U1, the same traffic and arithmetic on values of the state's own range)"""
        P = _lib.P
        code = PRECISION_CODES[self.precision]
        descs = (_lib.NvStageDesc * self.n)()
        for s, m in enumerate(self.members):
            d = descs[s]
            U = [t.data_ptr() for t in m.U]
            U1 = [m.stage[c].data_ptr() for c in range(4)]
            src, u0, dst = (U, U, U1) if stage == 1 else (U1, U, U)
            if scratch is not None:
                src, u0, dst = (U, U, scratch[s]) if stage == 1 else (U, U1, scratch[s])
            for c in range(4):
                d.u[c] = src[c]
                d.u0[c] = u0[c]
                d.out[c] = dst[c]
            d.xi = m.xi.data_ptr()
            d.stars = m.star_arr.data_ptr()
            d.jets = m.jet_arr.data_ptr()
            d.smax_out = (smax + 4 * s) if stage == 2 else None
            d.dt_dev = self.dt_f.data_ptr() + 4 * s
            d.nx, d.ny, d.nz = m.shape
            d.nstars = len(m.stars)
            d.wind = m.P["wind"]
            d.jet_cos = m.C["jet_cos"]
            d.jet_r = m.P["jet_r"]
            d.pulse = pulses[s]
            d.dt = 0.0
            d.dx = m.dx
            d.a, d.b = (0.0, 1.0) if stage == 1 else (0.5, 0.5)
            d.floor_ = m.P["floor"]
            d.cn2, d.ci2, d.vmax = m.cn2, m.ci2, 6.0

        def build(buf, nbytes):
            _lib.call("nv_hydro_stage_batch_plan", ctypes.cast(descs, P), self.n, code, self.solver, self.lim,
                      self.tile[stage], buf, nbytes)
        return _Plan("nv_hydro_stage_batch_plan_size", build, self.n, self.device)

    def _stage_plans(self, pulses):
        # the tile first, then the dispatch form for that tile
        if self.tune:
            from . import autotune
            self.tune = False
            self.tile = autotune.tune_batch(self, pulses)     # {stage: tile}
            self._stage_key = None
        if self.dispatch == "tune":
            from . import autotune
            choice = autotune.tune_batch_dispatch(self, pulses)     # {stage: (use the queue, workgroups per CU)}
            self.dispatch = {st: c[0] for st, c in choice.items()}
            self.queue_wgs = {st: c[1] for st, c in choice.items()}
        # everything a plan bakes in: the pulses, the tile, and per member its live wind, its star count and the
        # addresses of its star and jet tables (Sim.replace_stars may reallocate them)
        key = (tuple(pulses), tuple(self.tile.items())) + tuple(
            (m.P["wind"], len(m.stars), m.star_arr.data_ptr(), m.jet_arr.data_ptr()) for m in self.members)
        if key == self._stage_key:
            return
        self._plans_stage = [self._stage_plan(1, pulses), self._stage_plan(2, pulses, smax=self.smax.data_ptr())]
        self._stage_key = key
        self._graphs.clear()

    def _ray_dirs(self, nu, nphi):
        """device address of the direction table of an (nu, nphi) ray grid (nv_ray_dirs), built once per batch and
        shared by every job on that grid"""
        tab = self._dirs.get((nu, nphi))
        if tab is None:
            tab = torch.empty(4 * nu * nphi, dtype=torch.float32, device=self.device)
            _lib.call("nv_ray_dirs", nu, nphi, tab.data_ptr(), self._stream)
            self._dirs[(nu, nphi)] = tab
        return tab.data_ptr()

    def _ion_plans(self):
        """ray-cast jobs of every switched-on star of every member, and the fused ionize pass per member"""
        key = tuple((tuple(self._on(m)), tuple(round(m.q_eff(i), 12) for i in self._on(m))) for m in self.members)
        if key == self._ion_key:
            return
        P = _lib.P
        jobs = []
        ions = (_lib.NvIonDesc * self.n)()
        form = self.members[0].ion_form   # every member has the batch's precision
        for s, m in enumerate(self.members):
            on = self._on(m)
            nu, nphi = m.rfronts.shape[1:]
            Lx, Ly, Lz = m.bmax
            # per active star its front range (nv_ray_range_floats), which the casts write and the ionize pass
            # reads to skip stars that change no bit of a cell
            per = int(_lib.load().nv_ray_range_floats(nu, nphi))
            rng = self._ranges.get(s)
            if rng is None or rng.numel() < per * len(m.rfronts):
                rng = self._ranges[s] = torch.zeros(per * len(m.rfronts), dtype=torch.float32, device=self.device)
            for slot, i in enumerate(on):
                j = _lib.NvRayDesc()
                j.rho = m.U[0].data_ptr()
                j.nx, j.ny, j.nz = m.shape
                j.sx, j.sy, j.sz = m.stars[i][:3]
                j.Q = m.q_eff(i)
                j.dr = m.dr
                j.nsteps = m.nsteps
                j.dx = m.dx
                j.rcap = m.diag
                j.bx, j.by, j.bz = Lx, Ly, Lz
                j.rfront = m.rfronts[slot].data_ptr()
                j.nu, j.nphi = nu, nphi
                j.wide = 1 if m.long_rays == "fp64" else 0
                j.dirs = self._ray_dirs(nu, nphi)
                j.rf_range = rng.data_ptr() + 4 * per * slot if self.cull_stars and len(on) >= 2 else None
                # the member's 16 B/cell stage buffer is free while ionize runs: a brick plan builds the density
                # bricks there and the casts sample them
                j.bricks = m._brick_buf.data_ptr() if m.ray_bricks is not False else None
                jobs.append(j)
            xyz = np.array([m.stars[i][:3] for i in on], dtype=np.float32).reshape(-1, 3)
            if len(on):
                m.star_xyz[:len(on)].copy_(torch.from_numpy(xyz))
            d = ions[s]
            half = m.dfront.dtype == torch.float16
            d.rfronts = m.rfronts.data_ptr()
            d.nu, d.nphi = nu, nphi
            d.star_xyz = m.star_xyz.data_ptr()
            d.nstars = len(on)
            d.dx = m.dx
            d.width = 1.2 * m.dx
            d.xi = m.xi.data_ptr()
            d.dfront_h = m.dfront.data_ptr() if half else None
            d.dfront_f = None if half else m.dfront.data_ptr()
            d.nx, d.ny, d.nz = m.shape
            for c in range(4):
                d.state[c] = m.U[c].data_ptr()
            d.cn2, d.ci2 = m.cn2, m.ci2
            d.smax_out = self.smax.data_ptr() + 4 * s
            d.rf_ranges = rng.data_ptr() if self.cull_stars and len(on) >= 2 else None
        self._plan_rays = None
        if jobs:
            def ray_plan(use_bricks):
                descs = []
                for j in jobs:
                    d = _lib.NvRayDesc.from_buffer_copy(j)
                    if not use_bricks:
                        d.bricks = None
                    descs.append(d)
                arr = (_lib.NvRayDesc * len(descs))(*descs)

                def build_rays(buf, nbytes, arr=arr, n=len(descs)):
                    _lib.call("nv_cast_rays_batch_plan", ctypes.cast(arr, P), n, buf, nbytes)
                return _Plan("nv_cast_rays_batch_plan_size", build_rays, len(descs), self.device)

            use = self.ray_bricks if any(j.bricks for j in jobs) else False
            if use == "auto":
                from . import autotune
                plans = {False: ray_plan(False), True: ray_plan(True)}
                use = autotune.tune_batch_rays(self, plans)
                self._plan_rays = plans[use]
            else:
                self._plan_rays = ray_plan(bool(use))
            self.ray_bricks_used = bool(use)

        def build_ion(buf, nbytes):
            _lib.call("nv_ionize_apply_batch_plan", ctypes.cast(ions, P), self.n, form, buf, nbytes)
        self._plan_ion = _Plan("nv_ionize_apply_batch_plan_size", build_ion, self.n, self.device)
        self._ion_key = key
        self._graphs.clear()

    # ------------------------------------------------------------------ the step
    def _launch_ionize(self):
        strm = self._stream
        _lib.call("nv_fill_f32", self.smax.data_ptr(), _f(0.0), self.n, strm)
        if self._plan_rays is not None:
            tk = self._t0("cast_rays")
            _lib.call("nv_cast_rays_batch", *self._plan_rays.args, strm)
            self._t1(tk)
        tk = self._t0("ionize_apply")
        _lib.call("nv_ionize_apply_batch", *self._plan_ion.args, strm)
        self._t1(tk)

    def _launch_hydro(self):
        strm = self._stream
        for stage in (1, 2):
            tk = self._t0("hydro_stage")
            if self.dispatch[stage]:
                _lib.call("nv_hydro_stage_batch_queue_ex", *self._plans_stage[stage - 1].args,
                          self.queue_wgs[stage], strm)
            else:
                _lib.call("nv_hydro_stage_batch", *self._plans_stage[stage - 1].args, strm)
            self._t1(tk)

    def _device_step(self, ion):
        strm = self._stream
        if ion:
            self._launch_ionize()
        _lib.call("nv_cfl_dt", self.smax.data_ptr(), self.cfl_dx.data_ptr(), self.dt_cap.data_ptr(),
                  self.dt_f.data_ptr(), self.dt_d.data_ptr(), self.t_dev.data_ptr(), self.n, strm)
        self._launch_hydro()

    def _replay(self, ion):
        g = self._graphs.get(ion)
        if g is None:
            # warm the launch path once outside capture, then record the step
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self._device_step(ion)
            self._graphs[ion] = g
            g.replay()
            return
        g.replay()

    def step(self, k=1):
        """advance every member k steps (Sim._main_step order, sim.py:1099-1114)"""
        self._check_members()
        for m in self.members:
            m._set_star_weights()
        if self.device_clock_ok():
            self._stage_plans([self._pulse(m) for m in self.members])
            self._ion_plans()
            for _ in range(k):
                ion = self.it % self.ion_every == 0
                if self.graph and self.timer is None:
                    self._replay(ion)
                else:
                    self._device_step(ion)
                self.it += 1
            self._sync_clocks(k)
            return
        for _ in range(k):
            self._host_step()

    def _host_step(self):
        ion = self.it % self.ion_every == 0
        if ion:
            self._ion_plans()
            self._launch_ionize()
        for m in self.members:
            m._set_star_weights()
        smax = self.smax.cpu().numpy()             # the one readback per step (sim.py:1012 per member)
        dts, pulses = [], []
        for s, m in enumerate(self.members):
            dt = m.P["cfl"] * m.dx / float(smax[s])
            if m.dt_cap is not None:
                dt = min(dt, m.dt_cap)
            dts.append(dt)
            pulses.append(self._pulse(m))
        self.dt_f.copy_(torch.tensor(np.array(dts, dtype=np.float32)))
        self.smax.zero_()
        self._stage_plans(pulses)
        self._launch_hydro()
        for s, m in enumerate(self.members):
            m.t += dts[s]
            m.dt = dts[s]
            m.it += 1
            m.version += 1
            m._smax_ok = False             # the member's own cached CFL max predates this step
        self.it += 1
        self.t_dev.copy_(torch.tensor([m.t for m in self.members], dtype=torch.float64))

    def _sync_clocks(self, k):
        t = self.t_dev.cpu().numpy()
        dt = self.dt_d.cpu().numpy()
        for s, m in enumerate(self.members):
            m.t = float(t[s])
            m.dt = float(dt[s])
            m.it += k
            m.version += k
            m._smax_ok = False
