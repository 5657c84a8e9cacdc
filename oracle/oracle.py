"""CPU oracle of the simulation step. TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module, and only as the
checker or the timed CPU baseline; the product (nebulaviewer_amd) never imports it.

liboracle.so (oracle/nebula_oracle.cpp) restates the reference Warp kernels in C++ (float and double);
OracleSim restates the orchestration of the reference Sim step by step, kernel launch by kernel launch:
  ionize          sim.py:927-947  (xi.zero_, df.fill_(1e3), per star cast_rays + apply_front, smooth_xi,
                                   copy, to_half)
  star weights    sim.py:914-925
  _compute_dt     sim.py:1009-1012 (max_speed, dt = cfl dx / smax in double)
  jet pulse       sim.py:1108
  _hydro          sim.py:1014-1028 (stage 1 into the scratch, stage 2 in place)
  clocks          sim.py:1110-1114, version sim.py:1147
  fuv_field       sim.py:949-978
Scalars are computed in double and rounded to the kernel precision, as a Warp launch does.
"""
import ctypes
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        for s, R in (("f32", ctypes.c_float), ("f64", ctypes.c_double)):
            P, I = ctypes.c_void_p, ctypes.c_int
            getattr(_lib, f"or_hydro_stage_{s}").argtypes = [P] * 13 + [I, I, I, P, I, R, P] + [R] * 11 + [I, I, I]
            getattr(_lib, f"or_max_speed_{s}").argtypes = [P] * 5 + [ctypes.c_longlong, R, R, I]
            getattr(_lib, f"or_max_speed_{s}").restype = R
            getattr(_lib, f"or_cast_rays_{s}").argtypes = [P, I, I, I] + [R] * 5 + [I] + [R] * 5 + [P, I, I, I]
            getattr(_lib, f"or_apply_front_{s}").argtypes = [P, I, I] + [R] * 5 + [P, P, I, I, I, I]
            getattr(_lib, f"or_smooth_xi_{s}").argtypes = [P, P, I, I, I, I]
            getattr(_lib, f"or_cast_fuv_{s}").argtypes = [P, P, I, I, I] + [R] * 5 + [I] + [R] * 5 + [P, I, I, I, I]
            getattr(_lib, f"or_apply_fuv_{s}").argtypes = [P, I, I, I] + [R] * 6 + [P, I, I, I, I]
            getattr(_lib, f"or_sample_n_{s}").argtypes = [P, I, I, I, R, R, R, R]
            getattr(_lib, f"or_sample_n_{s}").restype = R
            getattr(_lib, f"or_box_range_{s}").argtypes = [R] * 9 + [P]
            getattr(_lib, f"or_face_flux_{s}").argtypes = [P, R, R, I, R, I, I, P]
            getattr(_lib, f"or_slope_{s}").argtypes = [R, R, I]
            getattr(_lib, f"or_slope_{s}").restype = R
        _lib.or_cast_rays_wide_f32.argtypes = _lib.or_cast_rays_f32.argtypes
        _lib.or_to_half.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_longlong, ctypes.c_int]
        for fn in ("or_sinf_det", "or_cosf_det", "or_expf_det"):
            getattr(_lib, fn).argtypes = [ctypes.c_float]
            getattr(_lib, fn).restype = ctypes.c_float
        _lib.or_atan2f_det.argtypes = [ctypes.c_float, ctypes.c_float]
        _lib.or_atan2f_det.restype = ctypes.c_float
        _lib.or_atan2f_det_n.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_longlong, ctypes.c_int]
    return _lib


class Codec:
    """The oracle's restatement of the consumers' kernels (oracle/codec_oracle.cpp)."""

    def __init__(self):
        self.lib = load()
        P, I, F, LL = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_longlong
        L = self.lib
        L.or_timeline_pack.argtypes = [P] * 14 + [LL]
        L.or_timeline_unpack.argtypes = [P] * 7 + [F] + [P] * 7 + [LL]
        L.or_shock_field.argtypes = [P] * 4 + [I, I, I, F, P]
        L.or_pack_volume.argtypes = [P] * 7 + [I, I, I, P, I, P]
        L.or_smooth_axis.argtypes = [P, P, I, I, I, I]

    def timeline_pack(self, rho, mx, my, mz, xi, df, g0):
        n = rho.size
        out = [np.empty(rho.shape, np.uint16) for _ in range(4)] + [np.empty(rho.shape, np.uint8)] + \
              [np.empty(rho.shape, np.uint16) for _ in range(2)]
        ins = [np.ascontiguousarray(a, np.float32) for a in (rho, mx, my, mz, xi, df, g0)]
        self.lib.or_timeline_pack(*[_ptr(a) for a in ins], *[_ptr(a) for a in out], n)
        return out

    def timeline_unpack(self, arrs, floor):
        shape = arrs[0].shape
        out = [np.empty(shape, np.float32) for _ in range(7)]
        self.lib.or_timeline_unpack(*[_ptr(a) for a in arrs], ctypes.c_float(floor), *[_ptr(a) for a in out],
                                    arrs[0].size)
        return out

    def shock_field(self, rho, mx, my, mz, dx):
        out = np.empty(rho.shape, np.float32)
        self.lib.or_shock_field(_ptr(rho), _ptr(mx), _ptr(my), _ptr(mz), *rho.shape, ctypes.c_float(dx), _ptr(out))
        return out

    def pack_volume(self, rho, mx, my, mz, xi, df, g0, pp, shock_on):
        """pp: 22 floats in the PackParams order; df, g0 as float32 values"""
        out = np.empty(rho.shape + (4,), np.uint16)
        ppa = np.ascontiguousarray(pp, np.float32)
        ins = [np.ascontiguousarray(a, np.float32) for a in (rho, mx, my, mz, xi, df, g0)]
        self.lib.or_pack_volume(*[_ptr(a) for a in ins], *rho.shape, _ptr(ppa), int(shock_on), _ptr(out))
        return out

    def drive(self, rho, m, F, G, blend, amp, dt, dx, dxs):
        """m (3 arrays) updated in place"""
        P, I, R = ctypes.c_void_p, ctypes.c_int, ctypes.c_float
        self.lib.or_drive.argtypes = [P] * 4 + [I, I, I] + [P] * 6 + [I, I, I] + [R] * 5
        self.lib.or_drive(_ptr(rho), *[_ptr(a) for a in m], *rho.shape, *[_ptr(a) for a in F], *[_ptr(a) for a in G],
                          *F[0].shape, R(blend), R(amp), R(dt), R(dx), R(dxs))

    def add_torus(self, rho, m, torus):
        self.lib.or_add_torus.argtypes = [ctypes.c_void_p] * 5 + [ctypes.c_longlong]
        self.lib.or_add_torus(_ptr(rho), *[_ptr(a) for a in m], _ptr(torus), rho.size)

    def mach_sums(self, rho, m):
        out = np.zeros(2)
        self.lib.or_mach_sums.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_longlong, ctypes.c_void_p]
        self.lib.or_mach_sums(_ptr(rho), *[_ptr(a) for a in m], rho.size, _ptr(out))
        return out

    def smooth_axis(self, src, axis):
        dst = np.empty_like(src)
        self.lib.or_smooth_axis(_ptr(src), _ptr(dst), axis, *src.shape[:3])
        return dst


def _ptr(a):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


class Kernels:
    """Thin typed wrappers of the oracle kernels for one precision."""

    def __init__(self, dtype=np.float32, threads=None):
        self.lib = load()
        self.dtype = np.dtype(dtype)
        self.s = "f32" if self.dtype == np.float32 else "f64"
        self.R = ctypes.c_float if self.s == "f32" else ctypes.c_double
        self.threads = threads or os.cpu_count() or 1

    def fn(self, name):
        return getattr(self.lib, f"or_{name}_{self.s}")

    def hydro_stage(self, U, xi, U0, out, shape, stars, nstars, wind, jets, jet_cos, jet_r, pulse, dt, dx, a, b,
                    floor, cn2, ci2, vmax, solver, lim):
        R = self.R
        self.fn("hydro_stage")(*[_ptr(x) for x in U], _ptr(xi), *[_ptr(x) for x in U0], *[_ptr(x) for x in out],
                               *shape, _ptr(stars), nstars, R(wind), _ptr(jets), R(jet_cos), R(jet_r), R(pulse),
                               R(dt), R(dx), R(a), R(b), R(floor), R(cn2), R(ci2), R(vmax), solver, lim,
                               self.threads)

    def max_speed(self, rho, mx, my, mz, xi, cn2, ci2):
        return float(self.fn("max_speed")(_ptr(rho), _ptr(mx), _ptr(my), _ptr(mz), _ptr(xi), rho.size,
                                          self.R(cn2), self.R(ci2), self.threads))

    def cast_rays(self, rho, star, Q, dr, nsteps, dx, rcap, bmax, rfront, wide=False):
        """wide: FP64 recombination column on rays longer than 512 cells (float kernels only)"""
        R = self.R
        if wide and self.s != "f32":
            raise ValueError("the FP64 long-ray column is a variant of the float kernel")
        self.fn("cast_rays_wide" if wide else "cast_rays")(_ptr(rho), *rho.shape, R(star[0]), R(star[1]), R(star[2]), R(Q), R(dr), nsteps, R(dx),
                             R(rcap), R(bmax[0]), R(bmax[1]), R(bmax[2]), _ptr(rfront), *rfront.shape, self.threads)

    def apply_front(self, rfront, star, dx, width, xi, df):
        R = self.R
        self.fn("apply_front")(_ptr(rfront), *rfront.shape, R(star[0]), R(star[1]), R(star[2]), R(dx), R(width),
                               _ptr(xi), _ptr(df), *xi.shape, self.threads)

    def smooth_xi(self, src, dst):
        self.fn("smooth_xi")(_ptr(src), _ptr(dst), *src.shape, self.threads)

    def cast_fuv(self, rho, xi, star, kappa, dr, nsteps, dx, binr, bmax, att):
        R = self.R
        self.fn("cast_fuv")(_ptr(rho), _ptr(xi), *rho.shape, R(star[0]), R(star[1]), R(star[2]), R(kappa), R(dr),
                            nsteps, R(dx), R(binr), R(bmax[0]), R(bmax[1]), R(bmax[2]), _ptr(att), *att.shape,
                            self.threads)

    def apply_fuv(self, att, star, lum, dx, binr, g0):
        R = self.R
        self.fn("apply_fuv")(_ptr(att), *att.shape, R(star[0]), R(star[1]), R(star[2]), R(lum), R(dx), R(binr),
                             _ptr(g0), *g0.shape, self.threads)

    def to_half(self, src):
        dst = np.empty(src.shape, np.uint16)
        self.lib.or_to_half(_ptr(np.ascontiguousarray(src, np.float32)), _ptr(dst), src.size, self.threads)
        return dst.view(np.float16)


class OracleSim:
    """The reference Sim step on the CPU (float32 = parity oracle, float64 = error-budget oracle).

    setup: nebulaviewer_amd.scenarios.Setup; consts: nebulaviewer_amd.scenarios.constants(setup, ...);
    rays = (nu, nphi); fuv = (nr, nu, nphi); long_rays = "fp64": the float oracle keeps the column of rays whose
    chord exceeds 512 cells in FP64 (or_cast_rays_wide_f32), "fp32": the reference's float column everywhere."""

    def __init__(self, setup, consts, rays, fuv, dtype=np.float32, threads=None, long_rays="fp64"):
        self.K = Kernels(dtype, threads)
        self.wide = long_rays == "fp64" and self.K.s == "f32"
        dt_ = self.K.dtype
        self.setup, self.C = setup, consts
        self.P = dict(setup.P)
        self.shape = tuple(setup.shape)
        self.dx = setup.dx
        self.rho = np.ascontiguousarray(setup.rho, dtype=dt_)
        z = lambda a: np.zeros(self.shape, dt_) if a is None else np.ascontiguousarray(a, dtype=dt_)
        self.mx, self.my, self.mz = z(setup.mx), z(setup.my), z(setup.mz)
        self.xi = np.zeros(self.shape, dt_)
        self.stage = np.zeros((4,) + self.shape, dt_)
        self.half = not self.P.get("fp32_fields") and dt_ == np.float32
        self.dfront = np.full(self.shape, 1e3, np.float16 if self.half else dt_)
        self.g0 = np.zeros(self.shape, np.float16 if self.half else dt_)
        self.stars = [tuple(float(v) for v in s) for s in setup.stars]
        self.kinds, self.jets, self.jkinds = list(setup.kinds), [tuple(j[:3]) for j in setup.jets], list(setup.jkinds)
        self.cap = max(len(self.stars), self.P["max_stars"])
        self.rfront = np.zeros(tuple(rays), dt_)
        self.att = np.zeros(tuple(fuv), dt_)
        self.t, self.it, self.dt, self.version = 0.0, 0, 0.0, 0
        # pre-turbulence phase (sim.py:786-789, 1030-1060); the float oracle only
        from nebulaviewer_amd import scenarios as S
        self.tp, self.ip, self.mach = 0.0, 0, None
        self.torus = None if setup.torus is None else np.ascontiguousarray(setup.torus, dtype=np.float32)
        self.stirring = S.Stirring(setup)
        self.drv = None
        if self.P.get("pre_t", 0.0) > 0:
            if dt_ != np.float32:
                raise ValueError("the pre-turbulence phase is restated for the float oracle")
            self.codec = Codec()
            self._set_star_weights()          # allocated; the pre-phase stages pass nstars = 0
            self.phase = "pre"
            self.drv = {"A": self.stirring.next(), "B": self.stirring.next(), "clock": 0.0}
        else:
            self.phase = "main"
            self._set_star_weights()
            self.ionize()

    def _start_main(self):
        if self.torus is not None:
            self.codec.add_torus(self.rho, [self.mx, self.my, self.mz], self.torus)
            self.torus = None
        self.phase = "main"
        self._set_star_weights()
        self.ionize()

    def _stir(self, dt, amp):
        d = self.drv
        blend = min(d["clock"] / self.P["pre_corr"], 1.0)
        self.codec.drive(self.rho, [self.mx, self.my, self.mz], d["A"], d["B"], blend, amp, dt, self.dx,
                         self.stirring.DXS)
        d["clock"] += dt
        if d["clock"] >= self.P["pre_corr"]:
            d["A"], d["B"], d["clock"] = d["B"], self.stirring.next(), 0.0

    def _pre_step(self):
        dt = self._compute_dt()
        self._hydro(dt, 0, 0.0, 0.0)
        self._stir(dt, self.P["pre_amp"])
        self.tp += dt
        self.ip += 1
        self.dt = dt
        if self.tp >= self.P["pre_t"]:
            a, b = self.codec.mach_sums(self.rho, [self.mx, self.my, self.mz])
            self.mach = math.sqrt(a / b) * self.P["ci_over_cn"]
            if self.P["drive_main"] <= 0:
                self.drv = None
            self._start_main()

    def q_of(self, i):
        s = self.stars[i]
        return s[3] * self.P["q_main"] if self.kinds[i] == "main" else s[3]

    def ramp(self, i):
        age = self.t - self.stars[i][4]
        if age < 0.0:
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

    def _set_star_weights(self):
        qm = self.P["q_main"]
        w = [(s[0], s[1], s[2], (self.q_eff(i) / qm) if self.t >= s[4] else 0.0) for i, s in enumerate(self.stars)]
        w += [(0.0, 0.0, 0.0, 0.0)] * (self.cap - len(w))
        # the reference uploads float32 vec4 tables (wp.vec4 arrays), also for its double restatement here
        self.star_arr = np.ascontiguousarray(np.array(w, dtype=np.float32).astype(self.K.dtype))
        jw = [(a[0], a[1], a[2], self.jet_of(i)) for i, a in enumerate(self.jets)]
        jw += [(0.0, 0.0, 1.0, 0.0)] * (self.cap - len(jw))
        self.jet_arr = np.ascontiguousarray(np.array(jw, dtype=np.float32).astype(self.K.dtype))

    def ionize(self):
        K, C = self.K, self.C
        df = self.stage[1]
        self.xi[...] = 0
        df[...] = 1e3
        for i, (sx, sy, sz, _, ton) in enumerate(self.stars):
            if self.t < ton:
                continue
            K.cast_rays(self.rho, (sx, sy, sz), self.q_eff(i), C["dr"], C["nsteps"], self.dx, C["diag"], C["bmax"],
                        self.rfront, wide=self.wide)
            K.apply_front(self.rfront, (sx, sy, sz), self.dx, 1.2 * self.dx, self.xi, df)
        K.smooth_xi(self.xi, self.stage[0])
        self.xi[...] = self.stage[0]
        self.dfront[...] = K.to_half(df) if self.half else df

    def fuv_field(self):
        K, C = self.K, self.C
        acc = self.stage[2]
        acc[...] = 0
        binr = C["diag"] / self.att.shape[0]
        for i, (sx, sy, sz, _, ton) in enumerate(self.stars):
            if self.t < ton:
                continue
            K.cast_fuv(self.rho, self.xi, (sx, sy, sz), self.P["kappa_fuv"], C["dr"], C["nsteps"], self.dx, binr,
                       C["bmax"], self.att)
            K.apply_fuv(self.att, (sx, sy, sz), self.q_eff(i) / self.P["q_main"], self.dx, binr, acc)
        self.g0[...] = K.to_half(acc) if self.half else acc

    def _compute_dt(self):
        smax = self.K.max_speed(self.rho, self.mx, self.my, self.mz, self.xi, self.C["cn2"], self.C["ci2"])
        return self.P["cfl"] * self.dx / smax

    def _hydro(self, dt, nstars, wind, pulse):
        P, C = self.P, self.C
        hll = P.get("solver", "rusanov") == "hll"
        lim = {"minmod": 0, "vanleer": 1, "mc": 2}[P.get("limiter", "minmod")] if hll else 0
        U = [self.rho, self.mx, self.my, self.mz]
        U1 = [self.stage[c] for c in range(4)]
        common = dict(shape=self.shape, stars=self.star_arr, nstars=nstars, wind=wind, jets=self.jet_arr,
                      jet_cos=C["jet_cos"], jet_r=P["jet_r"], pulse=pulse, dt=dt, dx=self.dx, floor=P["floor"],
                      cn2=C["cn2"], ci2=C["ci2"], vmax=6.0, solver=int(hll), lim=lim)
        self.K.hydro_stage(U, self.xi, U, U1, a=0.0, b=1.0, **common)
        self.K.hydro_stage(U1, self.xi, U, U, a=0.5, b=0.5, **common)

    def step(self, k=1):
        P = self.P
        for _ in range(k):
            if self.phase == "pre":
                self._pre_step()
                self.version += 1
                continue
            if self.it % self.C["ion_every"] == 0:
                self.ionize()
            self._set_star_weights()
            dt = self._compute_dt()
            pulse = max(0.0, 1.0 + P["jet_pulse"] * math.sin(2 * math.pi * self.t / P["jet_period"]))
            self._hydro(dt, len(self.stars), P["wind"], pulse)
            if P["drive_main"] > 0 and self.drv is not None:
                self._stir(dt, P["pre_amp"] * P["drive_main"])
            self.t += dt
            self.it += 1
            self.dt = dt
            self.version += 1
