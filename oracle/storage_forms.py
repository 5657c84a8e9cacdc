"""CPU measurement of reduced-storage forms of the conserved state (test infrastructure, not the product).

Each form keeps the arithmetic of the float32 oracle and rounds the conserved state to its storage encoding after
every RK stage (the stage's output, U1 and U^n+1, as a GPU stage would store it), and its density is compared with
the FP64 oracle of the same configuration (relative L1), at the grid sizes of tests/test_gpu_precision.py:

  fp16-vel     float32 rho, binary16 velocity m / rho (the form the library builds: nv_hydro_stage_h)
  log16-mtile  binary16 log2(rho) and binary16 momentum m / s_t, with s_t a power of two >= the largest |m_c| of
               the tile t = 32 consecutive z cells of one (x, y) row (log-density plus a per-tile momentum scale)
  logbf-mtile  the same in bfloat16 (8 significand bits)

    python -m oracle.storage_forms [config ...]      # configs: torus sheet wall torus_jet (default: all)

It prints one line per configuration and form; the figures are recorded in DESIGN.md. The oracle is built by
`make -C oracle` (or __graft_entry__.build()).
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path[:] = [p for p in sys.path if os.path.abspath(p or ".") != os.path.dirname(os.path.abspath(__file__))]
sys.path.insert(0, ROOT)

from nebulaviewer_amd import scenarios as S   # noqa: E402
from oracle.oracle import OracleSim            # noqa: E402

CONFIGS = {"torus": ("torus", 64, 100), "sheet": ("sheet", 48, 500), "wall": ("wall", 48, 300),
           "torus_jet": ("torus_jet", 48, 200)}
TILE = 32


def _bf16(x):
    """round float32 to bfloat16 (nearest, ties to even), returned as float32"""
    b = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32)


def _f16(x):
    return np.asarray(x, np.float32).astype(np.float16).astype(np.float32)


def round_state(U, form, floor):
    """the conserved state U = [rho, mx, my, mz] (float32, modified in place) as the form stores it"""
    rho, m = U[0], U[1:]
    if form == "fp16-vel":
        r = np.maximum(rho, np.float32(floor))
        for c in m:
            c[...] = rho * _f16(c / r)
        return
    q = _f16 if form == "log16-mtile" else _bf16
    rho[...] = np.exp2(q(np.log2(np.maximum(rho, np.float32(1e-30))))).astype(np.float32)
    nx, ny, nz = rho.shape
    nt = (nz + TILE - 1) // TILE
    pad = nt * TILE - nz
    amax = np.max(np.stack([np.abs(c) for c in m]), axis=0)
    amax = np.pad(amax, ((0, 0), (0, 0), (0, pad))).reshape(nx, ny, nt, TILE).max(axis=3)
    s = np.exp2(np.ceil(np.log2(np.maximum(amax, np.float32(1e-30))))).astype(np.float32)
    s = np.repeat(s, TILE, axis=2)[:, :, :nz]
    for c in m:
        c[...] = q(c / s) * s


class RoundedOracle(OracleSim):
    """the float32 oracle whose RK stages store their outputs in a reduced form"""

    def __init__(self, setup, consts, rays, fuv, form):
        self.form = form
        super().__init__(setup, consts, rays, fuv, np.float32, threads=os.cpu_count())
        round_state([self.rho, self.mx, self.my, self.mz], form, self.P["floor"])

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
        round_state(U1, self.form, P["floor"])
        self.K.hydro_stage(U1, self.xi, U, U, a=0.5, b=0.5, **common)
        round_state(U, self.form, P["floor"])


def rel_l1(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.abs(a - b).sum() / np.abs(b).sum())


def main(names):
    for name in names:
        scenario, n, steps = CONFIGS[name]
        setup = S.make_setup(scenario, n)
        C = S.constants(setup)
        rays, fuv = S.ray_grid(n), S.fuv_grid(n)
        ref = OracleSim(setup, C, rays, fuv, np.float64, threads=os.cpu_count())
        ref.step(steps)
        for form in ("fp16-vel", "log16-mtile", "logbf-mtile"):
            sim = RoundedOracle(S.make_setup(scenario, n), S.constants(setup), rays, fuv, form)
            sim.step(steps)
            errs = [rel_l1(getattr(sim, f), getattr(ref, f)) for f in ("rho", "mx", "my", "mz")]
            print(f"{name} {n}^3 {steps} steps  {form:12s} rel-L1 rho {errs[0]:.2e}  m {errs[1]:.2e} {errs[2]:.2e} "
                  f"{errs[3]:.2e}", flush=True)


if __name__ == "__main__":
    main(sys.argv[1:] or list(CONFIGS))
