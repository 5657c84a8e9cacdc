"""Parameters and seeded synthetic initial conditions for the simulation step.

Units are the reference's (sim.py:740-741): c_i = 1, c_n = 1/ci_over_cn = 1/12, density floor 2e-3
(nebula3d.py:591), |v| cap 6 (sim.py:1020), box [0, L] with dx = 1/res, outflow boundaries.

Parameters read by the step keep the names and defaults of nebula3d.main's argparse (nebula3d.py:580-639)
and sim.EXTRA_DEFAULTS (sim.py:75-76); the workloads select bake accuracy (Sim(fast=False): half-voxel rays,
ionization every P['ion_every'] steps, sim.py:743, 758) and the HLL + minmod solver.

The workloads are synthetic seeded scenes, one per benchmark configuration:
  torus  - 64^3 background n=1 plus a Gaussian torus of peak n=50 (major radius 0.2, minor 0.06) times a
           lognormal field (sigma 0.3, k^-4, seed 0); v = 0; one star at the centre with
           Q0 = 0.15^3/3 (uniform-n=1 Stromgren radius 0.15). 100 steps.
  sheet  - 256^3 slab |y - 1/2| < 0.1 of n=20 in n=1, lognormal sigma 0.5 (seed 1); solenoidal k^-4
           velocity at rms Mach 5 (|v| = 5 c_n); 4 stars in the sheet with Q ~ U[0.5, 2] Q0; wind 0.5;
           kappa_fuv 1. 500 steps.
  wall   - 512^3 half-space wall n=100 (tanh edge 4 cells) times lognormal sigma 0.4 (seed 2) plus 30 globules
           n=500, radius U[4, 12] cells; one star 0.4 above the face with Q = 0.4^3/3. 300 steps.
  torus_jet - the torus scene at 1024^3 (seed 3) with a +-z jet on the star. 200 steps.
  batch  - 64 independent members of uneven size (batch_setups): torus or sheet, N in {48..160}, swept Q and
           wind. 200 steps.
The random fields use the Gaussian-random-field recipe of nebula3d.grf3 (nebula3d.py:391-401): amplitude
k^(-slope/2) times complex normals from numpy.random.default_rng(seed), inverse real FFT, standardised; and
the lognormal map exp(sigma g - sigma^2/2) (nebula3d.py:427).
"""
import math
from dataclasses import dataclass, field

import numpy as np

Q0 = 0.15 ** 3 / 3.0

# parameters the step reads (nebula3d.py:580-639 defaults; sim.py:75-76 extras; workload choices)
STEP_DEFAULTS = dict(
    cfl=0.3, ci_over_cn=12.0, floor=2e-3, wind=0.0, ion_every=2, kappa_fuv=25.0,
    jet=0.0, jet_proto=0.0, jet_angle=10.0, jet_r=0.04, jet_pulse=0.6, jet_period=0.02,
    max_stars=32, q_main=Q0, q_ramp=0.0, solver="hll", limiter="minmod", fp32_fields=False, seed=0,
    # pre-turbulence phase (nebula3d.py:619-625): duration, stirring amplitude, band, correlation time; drive_main
    # keeps stirring in the main phase at drive_main x pre_amp
    pre_t=0.0, pre_amp=7.0, pre_kmax=4.0, pre_corr=0.06, drive_main=0.0,
)
# parameters that may change while the sim runs (sim.py:82-83); the others need a reset
LIVE = ("q_main", "wind", "jet", "jet_proto", "jet_pulse", "jet_period", "kappa_fuv", "drive_main", "pre_amp",
        "q_ramp")
LIMITERS = {"minmod": 0, "vanleer": 1, "mc": 2}
SOLVERS = {"rusanov": 0, "hll": 1}
BYTES_PER_CELL = 40   # rho, mx, my, mz, xi fp32 + stage 16 B + dfront, g0 fp16 (sim.py:94)


def grf3(shape, slope, rng):
    """Unit-variance Gaussian random field with power ~ k^-slope (the nebula3d.grf3 recipe)."""
    kx = np.fft.fftfreq(shape[0])[:, None, None]
    ky = np.fft.fftfreq(shape[1])[None, :, None]
    kz = np.fft.rfftfreq(shape[2])[None, None, :]
    k = np.sqrt(kx * kx + ky * ky + kz * kz)
    k[0, 0, 0] = 1.0
    amp = k ** (-slope / 2.0)
    amp[0, 0, 0] = 0.0
    del k
    coef = amp * (rng.standard_normal(amp.shape) + 1j * rng.standard_normal(amp.shape))
    del amp
    f = np.fft.irfftn(coef, s=shape, axes=(0, 1, 2)).astype(np.float32)
    del coef
    return (f - f.mean()) / f.std()


def lognormal(shape, sigma, slope, rng):
    g = grf3(shape, slope, rng)
    return np.exp(np.float32(sigma) * g - np.float32(0.5 * sigma * sigma)).astype(np.float32)


def solenoidal(shape, slope, rng):
    """Divergence-free Gaussian random vector field, power ~ k^-slope, unit rms |v| (float32 components)."""
    kx = np.fft.fftfreq(shape[0])[:, None, None]
    ky = np.fft.fftfreq(shape[1])[None, :, None]
    kz = np.fft.rfftfreq(shape[2])[None, None, :]
    k2 = kx * kx + ky * ky + kz * kz
    k2[0, 0, 0] = 1.0
    amp = k2 ** (-slope / 4.0)
    amp[0, 0, 0] = 0.0
    F = [amp * (rng.standard_normal(amp.shape) + 1j * rng.standard_normal(amp.shape)) for _ in range(3)]
    kd = (kx * F[0] + ky * F[1] + kz * F[2]) / k2
    F = [F[0] - kx * kd, F[1] - ky * kd, F[2] - kz * kd]
    v = [np.fft.irfftn(f, s=shape, axes=(0, 1, 2)) for f in F]
    rms = math.sqrt(sum(float((c * c).mean()) for c in v))
    return [np.ascontiguousarray(c / rms, dtype=np.float32) for c in v]


@dataclass
class Setup:
    """Everything Sim.reset needs: grid, parameters, initial fields and the star list.

    stars: [(x, y, z, qf, t_on)] with Q = qf * q_main for kind 'main', Q = qf for kind 'abs' (sim.py:710-712);
    jets: [(ax, ay, az)] unit axes, jkinds 'jet' (strength P['jet']), 'proto' or 'none' (sim.py:634-636).
    """
    name: str
    shape: tuple
    res: int
    P: dict
    rho: np.ndarray
    mx: np.ndarray = None
    my: np.ndarray = None
    mz: np.ndarray = None
    torus: np.ndarray = None     # density added at rest when the pre-turbulence phase ends (sim.py:806-812)
    stars: list = field(default_factory=list)
    kinds: list = field(default_factory=list)
    jets: list = field(default_factory=list)
    jkinds: list = field(default_factory=list)
    steps: int = 100

    @property
    def dx(self):
        return 1.0 / self.res

    @property
    def L(self):
        return tuple(n * self.dx for n in self.shape)


def make_params(overrides=None):
    p = dict(STEP_DEFAULTS)
    if overrides:
        p.update(overrides)
    return p


def stirring_field(shape, rng, kmin=1.0, kmax=6.0):
    """A unit-RMS divergence-free random vector field on a small grid, band-limited to kmin <= |k| <= kmax (k in
    grid wavenumbers), amplitude 1/|k|: the large-scale stirring of the pre-turbulence phase. The steps are those
    of nebula3d.solenoidal (nebula3d.py:514-524): complex normal modes, the divergence projected out, the real
    part of the inverse FFT, normalised to unit RMS speed."""
    k = np.meshgrid(*[np.fft.fftfreq(n) * n for n in shape], indexing="ij")
    kk = np.sqrt(k[0] ** 2 + k[1] ** 2 + k[2] ** 2)
    band = (kk >= kmin) & (kk <= kmax)
    amp = np.zeros(shape)
    amp[band] = 1.0 / kk[band]
    modes = [amp * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) for _ in range(3)]
    k2 = np.maximum(kk * kk, 1e-12)
    div = (k[0] * modes[0] + k[1] * modes[1] + k[2] * modes[2]) / k2
    v = [np.fft.ifftn(modes[c] - k[c] * div).real for c in range(3)]
    rms = math.sqrt(sum(float((c * c).mean()) for c in v))
    return [np.ascontiguousarray(c / rms, dtype=np.float32) for c in v]


class Stirring:
    """The sequence of stirring fields of one run, from its own seeded stream, so that every implementation of
    the step draws the same fields. Grid: 24 cells per domain unit plus a guard cell each side (the reference's
    fixed-resolution forcing, sim.py:761-764)."""

    DXS = 1.0 / 24.0

    def __init__(self, setup):
        self.shape = tuple(int(math.ceil(L_ / self.DXS)) + 2 for L_ in setup.L)
        self.kmax = setup.P["pre_kmax"]
        self.rng = np.random.default_rng(int(setup.P.get("seed", 0)) + 7919)

    def next(self):
        return stirring_field(self.shape, self.rng, 1.0, self.kmax)


def _axes(n, dx):
    x = ((np.arange(n) + 0.5) * dx).astype(np.float32)
    return x[:, None, None], x[None, :, None], x[None, None, :]


def torus(n=64, seed=0, jet=False, params=None):
    """Background n=1 + Gaussian torus (peak 50, R=0.2, a=0.06, axis z) times lognormal(0.3); one central star."""
    P = make_params(dict(seed=seed, q_main=Q0))
    if jet:
        # a +-z jet on the star (the reference's cone acceleration, nebula3d.py:126-139): strength 20 c_n, pulse
        # period of about 50 steps. A step is cfl dx / smax; once the jet runs, the capped flow along its axis
        # (vmax = 6) plus the ionized sound speed (c_i = 1) sets smax ~ 7 (measured: mean 6.6, rising to 7.6, over
        # the first 400 steps at 256^3), so the period is 50 cfl dx / 7
        P.update(jet=20.0 / P["ci_over_cn"], jet_pulse=0.7, jet_period=50 * P["cfl"] / n / 7.0)
    if params:
        P.update(params)
    dx = 1.0 / n
    X, Y, Z = _axes(n, dx)
    rc = np.sqrt((X - 0.5) ** 2 + (Y - 0.5) ** 2)
    h = Z - np.float32(0.5)
    tor = 50.0 * np.exp(-((rc - 0.2) ** 2 + h * h) / np.float32(2 * 0.06 ** 2))
    rng = np.random.default_rng(seed)
    ln = lognormal((n, n, n), 0.3, 4.0, rng)
    torus_rho = None
    if P["pre_t"] > 0:
        # with a turbulent phase, the torus forms when it ends, at rest (the reference's bipolar scene)
        rho = np.maximum(ln, np.float32(P["floor"])).astype(np.float32)
        torus_rho = (tor.astype(np.float32) * ln).astype(np.float32)
    else:
        rho = (1.0 + tor).astype(np.float32) * ln
        rho = np.maximum(rho, np.float32(P["floor"])).astype(np.float32)
    stars = [(0.5, 0.5, 0.5, 1.0, 0.0)]
    jets = [(0.0, 0.0, 1.0)]
    return Setup("torus_jet" if jet else "torus", (n, n, n), n, P, rho, torus=torus_rho, stars=stars, kinds=["main"],
                 jets=jets, jkinds=["jet" if jet else "none"], steps=200 if jet else 100)


def sheet(n=256, seed=1, params=None):
    """Slab |y-1/2|<0.1 of n=20 in n=1 times lognormal(0.5); solenoidal Mach-5 velocity; 4 stars in the sheet."""
    P = make_params(dict(seed=seed, q_main=Q0, wind=0.5, kappa_fuv=1.0))
    if params:
        P.update(params)
    dx = 1.0 / n
    X, Y, Z = _axes(n, dx)
    rng = np.random.default_rng(seed)
    base = np.where(np.abs(Y - np.float32(0.5)) < np.float32(0.1), np.float32(20.0), np.float32(1.0))
    rho = (base * lognormal((n, n, n), 0.5, 4.0, rng)).astype(np.float32)
    rho = np.maximum(rho, np.float32(P["floor"])).astype(np.float32)
    v = solenoidal((n, n, n), 4.0, rng)
    vamp = np.float32(5.0 / P["ci_over_cn"])
    mx, my, mz = (rho * (vamp * c) for c in v)
    stars = []
    for _ in range(4):
        x, z = rng.uniform(0.15, 0.85), rng.uniform(0.15, 0.85)
        y = 0.5 + rng.uniform(-0.04, 0.04)
        stars.append((float(x), float(y), float(z), float(rng.uniform(0.5, 2.0)), 0.0))
    jets = [(0.0, 0.0, 1.0)] * 4
    return Setup("sheet", (n, n, n), n, P, rho, mx.astype(np.float32), my.astype(np.float32),
                 mz.astype(np.float32), stars=stars, kinds=["main"] * 4, jets=jets, jkinds=["none"] * 4, steps=500)


def wall(n=512, seed=2, params=None):
    """Half-space wall n=100 (tanh edge of 4 cells at z=0.45) times lognormal(0.4), 30 globules of n=500 near
    the face, one star at (0.5, 0.5, 0.85): 0.4 above the face."""
    P = make_params(dict(seed=seed, q_main=0.4 ** 3 / 3.0))
    if params:
        P.update(params)
    dx = 1.0 / n
    X, Y, Z = _axes(n, dx)
    zf = 0.45
    rng = np.random.default_rng(seed)
    base = (1.0 + 99.0 * 0.5 * (1.0 - np.tanh((Z - np.float32(zf)) / np.float32(4 * dx)))).astype(np.float32)
    rho = base * lognormal((n, n, n), 0.4, 4.0, rng)
    x1 = X[:, 0, 0]
    for _ in range(30):
        gx, gy = rng.uniform(0.05, 0.95), rng.uniform(0.05, 0.95)
        gz = zf + rng.uniform(-0.08, 0.03)
        R = rng.uniform(4.0, 12.0) * dx
        h = int(R / dx) + 3
        c = [min(max(int(v / dx), 0), n - 1) for v in (gx, gy, gz)]
        sl = tuple(slice(max(ci - h, 0), min(ci + h + 1, n)) for ci in c)
        rr = np.sqrt((x1[sl[0], None, None] - gx) ** 2 + (x1[None, sl[1], None] - gy) ** 2
                     + (x1[None, None, sl[2]] - gz) ** 2)
        glob = (500.0 * 0.5 * (1.0 - np.tanh((rr - R) / dx))).astype(np.float32)
        rho[sl] = np.maximum(rho[sl], glob)
    rho = np.maximum(rho, np.float32(P["floor"])).astype(np.float32)
    stars = [(0.5, 0.5, zf + 0.4, 1.0, 0.0)]
    return Setup("wall", (n, n, n), n, P, rho, stars=stars, kinds=["main"], jets=[(0.0, 0.0, 1.0)],
                 jkinds=["none"], steps=300)


SCENARIOS = {"torus": torus, "sheet": sheet, "wall": wall,
             "torus_jet": lambda n=1024, seed=3, params=None: torus(n, seed, jet=True, params=params)}

# the benchmark configurations: scenario, grid, seed
CONFIGS = {
    0: ("torus", 64, 0),
    1: ("sheet", 256, 1),
    2: ("wall", 512, 2),
    3: ("torus_jet", 1024, 3),
}


# the uneven batch: many independent nebulae (sweeps over luminosity, wind and seed), member i from seed
# base + i: grid N in {48, 64, 96, 128, 160} (E[N^3] = 1.49 M cells), torus for even seeds and sheet for odd,
# Q_main ~ U[0.3, 3] Q0, wind ~ U[0, 1]; 200 steps
BATCH_SIZES = (48, 64, 96, 128, 160)


def batch_setup(i, base_seed=100, sizes=BATCH_SIZES):
    """member i of the uneven batch: seed base_seed + i draws its size, Q and wind; torus for even seeds, sheet for
    odd ones"""
    seed = base_seed + i
    rng = np.random.default_rng(seed)
    n = int(rng.choice(sizes))
    params = dict(q_main=float(Q0 * rng.uniform(0.3, 3.0)), wind=float(rng.uniform(0.0, 1.0)))
    st = make_setup("torus" if seed % 2 == 0 else "sheet", n, seed, params)
    st.steps = 200
    return st


def batch_setups(count=64, base_seed=100, sizes=BATCH_SIZES):
    return [batch_setup(i, base_seed, sizes) for i in range(count)]


def make_setup(scenario, n=None, seed=None, params=None):
    fn = SCENARIOS[scenario]
    kw = {}
    if n is not None:
        kw["n"] = n
    if seed is not None:
        kw["seed"] = seed
    if params:
        kw["params"] = params
    return fn(**kw)


def ray_grid(nz, fast=False):
    """(nu, nphi) of the ionizing-ray grid (sim.py:753-754)."""
    nu = 512 if nz <= 256 else 1024
    return nu, 2 * nu


def fuv_grid(nz):
    """(nr, nu, nphi) of the FUV attenuation grid (sim.py:755-756)."""
    return (128, 128, 256) if nz <= 192 else ((160, 160, 320) if nz <= 256 else (256, 256, 512))


def constants(setup, fast=False, ion_every=None):
    """Derived constants of Sim.reset (sim.py:736-759), computed in double as the reference does."""
    P = setup.P
    Lx, Ly, Lz = setup.L
    dx = setup.dx
    diag = math.sqrt(Lx ** 2 + Ly ** 2 + Lz ** 2) + 0.6
    dr = (1.0 if fast else 0.5) * dx
    return dict(
        cn2=(1.0 / P["ci_over_cn"]) ** 2, ci2=1.0, diag=diag, dr=dr, nsteps=int(diag / dr) + 1,
        cap=max(len(setup.stars), P["max_stars"]), bmax=(Lx, Ly, Lz),
        ion_every=ion_every or (max(P["ion_every"], 3) if fast else P["ion_every"]),
        fuv_every=(8 if setup.shape[2] <= 160 else 16) if fast else 8,
        jet_cos=math.cos(math.radians(P["jet_angle"])), width=1.2 * dx,
    )
