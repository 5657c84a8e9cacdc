"""Closed forms that pin the oracle's hydro source terms (nebula3d.py:113-147, sim.py:463-494) and the jet pulse
(sim.py:1108).

On a uniform state every face of the grid carries the same flux (clamped ghosts included), so the flux divergence
of a stage is exactly zero and a stage with a = 0, b = 1 leaves U + dt * (source terms) behind:
  - wind on ionized gas (xi > 0.01): dm = rho wind xi sum_s w_s rel_s / (|rel_s|^2 + 4e-4)^(3/2);
  - jets: dm = rho sum_s ax_s sign(cos) j_s pulse / (r_s^2 + 4e-4) inside the cone r < jet_r, r > 1e-6,
    |cos| > jet_cos of a switched-on star, nothing outside it;
  - the speed cap: a uniform flow faster than vmax leaves the stage at vmax in the same direction.
The expected values are evaluated here in float64 with numpy over the whole grid, independently of the oracle's
per-cell loop; each test is sensitive to the sign of rel, the softening 4e-4, the exponent 3/2 and the cone test.
"""
import math

import numpy as np
import pytest

RHO0, C2N, C2I = 2.5, 1.0 / 144.0, 1.0


def run_stage(K, shape, xi, stars, jets=None, wind=0.0, pulse=1.0, jet_cos=0.9, jet_r=0.3, dt=1e-3, vel=None,
              rho=RHO0, vmax=6.0):
    d = K.dtype.type
    U = [np.full(shape, rho, K.dtype)] + [np.zeros(shape, K.dtype) for _ in range(3)]
    if vel is not None:
        for c in range(3):
            U[c + 1][...] = rho * vel[c]
    xi = np.ascontiguousarray(np.broadcast_to(np.asarray(xi, K.dtype), shape))
    st = np.ascontiguousarray(np.asarray(stars, np.float32).reshape(-1, 4).astype(K.dtype))
    if jets is None:
        jets = [(0.0, 0.0, 1.0, 0.0)] * len(st)
    jt = np.ascontiguousarray(np.asarray(jets, np.float32).reshape(-1, 4).astype(K.dtype))
    out = [np.zeros(shape, K.dtype) for _ in range(4)]
    K.hydro_stage(U, xi, U, out, shape, st, len(st), d(wind), jt, d(jet_cos), d(jet_r), d(pulse), d(dt),
                  d(1.0 / shape[2]), d(0.0), d(1.0), d(2e-3), d(C2N), d(C2I), d(vmax), 1, 0)
    return U, out


def centres(shape):
    dx = 1.0 / shape[2]
    i, j, k = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in shape], indexing="ij")
    return np.stack([(i + 0.5) * dx, (j + 0.5) * dx, (k + 0.5) * dx])


def star_f64(s):
    """a star table entry as the kernels see it (float32 vec4)"""
    return np.asarray(s, np.float32).astype(np.float64)


# ------------------------------------------------------------------ wind (nebula3d.py:113-125, sim.py:463-473)
def wind_expected(shape, stars, wind, xi, rho=RHO0, dt=1e-3):
    p = centres(shape)
    acc = np.zeros((3,) + shape)
    for s in stars:
        x, y, z, w = star_f64(s)
        if w <= 0:
            continue
        rel = p - np.array([x, y, z])[:, None, None, None]
        r2 = (rel ** 2).sum(0) + 4e-4
        acc += rel * (w / r2 ** 1.5)
    xi = np.broadcast_to(np.asarray(xi, np.float64), shape)
    on = xi > 0.01
    return np.where(on, dt * rho * wind * xi * acc, 0.0)


STARS = [(0.31, 0.52, 0.47, 1.0), (0.70, 0.25, 0.66, 0.35), (0.5, 0.5, 0.5, 0.0)]   # the last one not yet on


@pytest.mark.parametrize("wind", [0.2, 1.0])
def test_wind_closed_form(okern64, wind):
    shape = (12, 14, 16)
    U, out = run_stage(okern64, shape, 1.0, STARS, wind=wind)
    np.testing.assert_array_equal(out[0], U[0])          # no source on the density
    expect = wind_expected(shape, STARS, wind, 1.0)
    for c in range(3):
        np.testing.assert_allclose(out[c + 1], expect[c], rtol=1e-12, atol=1e-18)
    # the term is not small against its own rounding: a dropped softening or exponent would show
    assert np.abs(expect).max() > 1e-3


def test_wind_needs_ionized_gas(okern64):
    # xi > 0.01 switches the term on (sim.py:466), scaled by xi; at xi <= 0.01 the momentum stays zero
    shape = (8, 10, 12)
    rng = np.random.default_rng(5)
    xi = rng.choice([0.0, 0.005, 0.01, 0.0100001, 0.4, 1.0], size=shape)
    U, out = run_stage(okern64, shape, xi, STARS, wind=0.6)
    # the flux divergence is no longer zero where c^2 varies, but it is the same with and without the wind:
    # the difference of the two stages is the wind term alone (to the rounding of U + dt dU)
    expect = wind_expected(shape, STARS, 0.6, xi)
    U2, out2 = run_stage(okern64, shape, xi, STARS, wind=0.0)
    for c in range(3):
        np.testing.assert_allclose(out[c + 1] - out2[c + 1], expect[c], rtol=1e-9, atol=1e-15)
    off = xi <= 0.01
    for c in range(3):
        np.testing.assert_array_equal(out[c + 1][off], out2[c + 1][off])


def test_wind_points_away_from_the_star(okern32):
    # one star in the middle of the box: every cell's momentum points radially outward
    shape = (16, 16, 16)
    star = [(0.5 + 1 / 64, 0.5 - 1 / 128, 0.5 + 3 / 256, 1.0)]
    U, out = run_stage(okern32, shape, 1.0, star, wind=1.0)
    p = centres(shape)
    rel = p - np.array(star[0][:3])[:, None, None, None]
    m = np.stack([out[1], out[2], out[3]]).astype(np.float64)
    cos = (m * rel).sum(0) / (np.linalg.norm(m, axis=0) * np.linalg.norm(rel, axis=0))
    assert cos.min() > 1 - 1e-5
    expect = wind_expected(shape, star, 1.0, 1.0)
    for c in range(3):
        np.testing.assert_allclose(out[c + 1], expect[c], rtol=2e-6, atol=1e-9)


# ------------------------------------------------------------------ jets (nebula3d.py:126-139, sim.py:474-485)
def jet_expected(shape, stars, jets, pulse, jet_cos, jet_r, rho=RHO0, dt=1e-3):
    p = centres(shape)
    acc = np.zeros((3,) + shape)
    for s, j in zip(stars, jets):
        x, y, z, w = star_f64(s)
        ax = np.asarray(j, np.float32).astype(np.float64)
        if ax[3] <= 0 or w <= 0:
            continue
        rel = p - np.array([x, y, z])[:, None, None, None]
        r = np.sqrt((rel ** 2).sum(0))
        cth = (rel * ax[:3, None, None, None]).sum(0) / np.maximum(r, 1e-30)
        inside = (r < jet_r) & (r > 1e-6) & (np.abs(cth) > jet_cos)
        f = np.where(inside, np.sign(cth) * ax[3] * pulse / (r * r + 4e-4), 0.0)
        acc += ax[:3, None, None, None] * f
        # no cell sits on the cone's edge within rounding (the closed form would not pin such a cell)
        assert np.abs(np.abs(cth) - jet_cos).min() > 1e-6 and np.abs(r - jet_r).min() > 1e-6
    return dt * rho * acc


JSTARS = [(0.5 + 0.013, 0.5 - 0.007, 0.5 + 0.011, 1.0), (0.27, 0.71, 0.33, 0.6), (0.8, 0.2, 0.8, 0.0)]
JETS = [(0.0, 0.0, 1.0, 2.0), (0.6, 0.0, 0.8, 0.7), (1.0, 0.0, 0.0, 5.0)]   # the third star is off: no jet


@pytest.mark.parametrize("pulse", [1.0, 0.35])
def test_jet_closed_form(okern64, pulse):
    shape = (16, 16, 16)
    jet_cos, jet_r = 0.9, 0.3
    U, out = run_stage(okern64, shape, 0.0, JSTARS, JETS, wind=0.0, pulse=pulse, jet_cos=jet_cos, jet_r=jet_r)
    np.testing.assert_array_equal(out[0], U[0])
    expect = jet_expected(shape, JSTARS, JETS, pulse, jet_cos, jet_r)
    for c in range(3):
        np.testing.assert_allclose(out[c + 1], expect[c], rtol=1e-12, atol=1e-18)
    hit = np.abs(expect).sum(0) > 0
    # both cones of the first star are populated, and most of the box is outside every cone
    assert 20 < hit.sum() < 0.2 * hit.size
    assert (expect[2][hit] > 0).any() and (expect[2][hit] < 0).any()


def test_jet_sign_follows_the_cone(okern32):
    # along +axis the momentum points along +axis, along -axis along -axis (wp.where(c > 0, 1, -1))
    shape = (16, 16, 16)
    star = [(0.5, 0.5, 0.5 + 1 / 64, 1.0)]
    jet = [(0.0, 0.0, 1.0, 1.0)]
    U, out = run_stage(okern32, shape, 0.0, star, jet, jet_cos=0.95, jet_r=0.4)
    mz = out[3].astype(np.float64)
    p = centres(shape)
    above = p[2] > star[0][2]
    assert (mz[above] >= 0).all() and (mz[~above] <= 0).all()
    assert (mz[above] > 0).sum() > 4 and (mz[~above] < 0).sum() > 4
    np.testing.assert_array_equal(out[1], 0)
    np.testing.assert_array_equal(out[2], 0)


def test_jet_pulse_of_the_step():
    # OracleSim passes pulse = max(0, 1 + jet_pulse sin(2 pi t / jet_period)) (sim.py:1108) to both stages
    from oracle.oracle import OracleSim
    captured = []

    class Probe(OracleSim):
        def __init__(self):
            self.P = dict(jet_pulse=0.0, jet_period=0.2, drive_main=0.0, wind=0.0)
            self.C = dict(ion_every=10 ** 9)
            self.it, self.t, self.version, self.phase, self.drv = 1, 0.0, 0, "main", None
            self.stars = []

        def _set_star_weights(self):
            pass

        def _compute_dt(self):
            return 1e-3

        def _hydro(self, dt, nstars, wind, pulse):
            captured.append(pulse)

    s = Probe()
    cases = [(0.0, 0.6, 1.0), (0.05, 0.6, 1.6), (0.15, 0.6, 0.4), (0.15, 1.5, 0.0), (0.1, 0.6, 1.0),
             (0.025, 0.6, 1.0 + 0.6 * math.sqrt(0.5))]
    for t, jp, expect in cases:
        s.t, s.P["jet_pulse"] = t, jp
        s.step(1)
        assert captured[-1] == pytest.approx(expect, abs=1e-12), (t, jp)


# ------------------------------------------------------------------ the speed cap (nebula3d.py:144-147)
@pytest.mark.parametrize("speed,direction", [(9.0, (0.6, 0.0, 0.8)), (7.5, (-0.48, 0.6, -0.64)),
                                             (6.5, (0.0, -1.0, 0.0))])
def test_speed_cap(okern64, speed, direction):
    shape = (6, 7, 8)
    vel = [speed * c for c in direction]
    U, out = run_stage(okern64, shape, 0.3, [(0.5, 0.5, 0.5, 0.0)], wind=0.0, vel=vel)
    np.testing.assert_array_equal(out[0], U[0])
    v = np.stack([out[c + 1] / out[0] for c in range(3)])
    np.testing.assert_allclose(np.linalg.norm(v, axis=0), 6.0, rtol=1e-13)
    for c in range(3):
        np.testing.assert_allclose(v[c], 6.0 * direction[c], rtol=1e-13, atol=1e-13)


def test_speed_below_cap_untouched(okern32):
    shape = (6, 7, 8)
    vel = (3.4, -2.9, 3.6)                       # |v| = 5.74 < 6
    U, out = run_stage(okern32, shape, 0.3, [(0.5, 0.5, 0.5, 0.0)], wind=0.0, vel=vel)
    for c in range(4):
        np.testing.assert_array_equal(out[c], U[c])
