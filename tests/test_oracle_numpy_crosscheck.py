"""The C++ oracle against the independent NumPy restatement (tests/oracle_numpy.py), bit for bit, in float32
and float64, on small random non-cubic grids with every solver, limiter and source term switched on."""
import math

import numpy as np
import pytest

import oracle_numpy as ON


def random_state(shape, dtype, seed, rho_lo=0.001, vel=1.0):
    rng = np.random.default_rng(seed)
    rho = rng.lognormal(0.0, 1.0, shape).clip(rho_lo, None).astype(dtype)
    m = [(rho * rng.normal(0, vel, shape)).astype(dtype) for _ in range(3)]
    xi = rng.uniform(0, 1, shape).astype(dtype)
    xi[rng.uniform(size=shape) < 0.3] = 0
    # a few superluminal cells (|v| > vmax = 6) and sub-floor densities, so the cap and the floor both act
    fast = rng.uniform(size=shape) < 0.05
    m[0][fast] = (rho[fast] * 9.0).astype(dtype)
    rho[rng.uniform(size=shape) < 0.03] = dtype(1e-4)
    return [rho, *m], xi


def star_tables(dtype, shape, dx):
    L = np.array(shape) * dx
    stars = np.array([[0.3 * L[0], 0.6 * L[1], 0.5 * L[2], 1.0], [0.7 * L[0], 0.2 * L[1], 0.4 * L[2], 0.5],
                      [0.5 * L[0], 0.5 * L[1], 0.5 * L[2], 0.0]], np.float32).astype(dtype)
    ax = np.array([0.3, -0.4, 0.866], np.float64)
    ax /= np.linalg.norm(ax)
    jets = np.array([[*ax, 2.0], [0, 0, 1, 0.0], [1, 0, 0, 3.0]], np.float32).astype(dtype)
    return np.ascontiguousarray(stars), np.ascontiguousarray(jets)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("solver,lim", [(1, 0), (1, 1), (1, 2), (0, 0)])
@pytest.mark.parametrize("stage2", [False, True])
def test_hydro_stage_matches_numpy(okern32, okern64, dtype, solver, lim, stage2):
    K = okern32 if dtype == np.float32 else okern64
    d = np.dtype(dtype).type
    shape = (12, 9, 14)
    dx = d(1.0 / 14)
    U, xi = random_state(shape, dtype, seed=10 + solver * 3 + lim)
    stars, jets = star_tables(dtype, shape, float(dx))
    if stage2:
        U0, _ = random_state(shape, dtype, seed=99)
        a, b = d(0.5), d(0.5)
    else:
        U0, a, b = U, d(0.0), d(1.0)
    # jet_r covers most of the box so the cone test and both signs are exercised
    args = dict(nstars=3, wind=d(0.7), jet_cos=d(math.cos(math.radians(30))), jet_r=d(0.6), pulse=d(1.3),
                dt=d(0.002), dx=dx, a=a, b=b, floor=d(2e-3), cn2=d((1 / 12) ** 2), ci2=d(1.0), vmax=d(6.0),
                solver=solver, lim=lim)
    out = [np.zeros(shape, dtype) for _ in range(4)]
    K.hydro_stage(U, xi, U0, out, shape, stars, jets=jets, **args)
    ref = ON.hydro_stage(U, xi, U0, stars, jets=jets, **args)
    for c in range(4):
        np.testing.assert_array_equal(out[c], ref[c], err_msg=f"field {c}")
    # the velocity cap and the density floor were both active in this state
    assert (np.sqrt(sum((o / out[0]) ** 2 for o in out[1:])) > 5.99).any()
    assert stage2 or (out[0] == d(2e-3)).any()


def test_hydro_stage_in_place_equals_copy(okern32):
    # stage 2 writes over U^n (sim.py:1026-1028): same bits as writing to a separate array
    shape = (10, 11, 12)
    d = np.float32
    U, xi = random_state(shape, d, 5)
    U1, _ = random_state(shape, d, 6)
    stars, jets = star_tables(d, shape, 1 / 12)
    args = dict(nstars=3, wind=d(0.5), jet_cos=d(0.9), jet_r=d(0.3), pulse=d(1.0), dt=d(0.001), dx=d(1 / 12),
                a=d(0.5), b=d(0.5), floor=d(2e-3), cn2=d(1 / 144), ci2=d(1), vmax=d(6), solver=1, lim=0)
    sep = [np.zeros(shape, d) for _ in range(4)]
    okern32.hydro_stage(U1, xi, U, sep, shape, stars, jets=jets, **args)
    inpl = [u.copy() for u in U]
    okern32.hydro_stage(U1, xi, inpl, inpl, shape, stars, jets=jets, **args)
    for c in range(4):
        np.testing.assert_array_equal(sep[c], inpl[c])


def det_sincos(lib):
    return lambda v: (lib.or_sinf_det(float(v)), lib.or_cosf_det(float(v)))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_cast_rays_matches_numpy(okern32, okern64, dtype):
    K = okern32 if dtype == np.float32 else okern64
    d = np.dtype(dtype).type
    shape = (16, 14, 18)
    dx = 1.0 / 18
    rng = np.random.default_rng(7)
    rho = rng.lognormal(0, 1, shape).astype(dtype)
    L = tuple(n * dx for n in shape)
    diag = math.sqrt(sum(v * v for v in L)) + 0.6
    dr = 0.5 * dx
    nsteps = int(diag / dr) + 1
    nu, nphi = 8, 16
    sincos = det_sincos(K.lib) if dtype == np.float32 else (lambda v: (math.sin(v), math.cos(v)))
    for star, Q in (((0.4, 0.35, 0.5), 0.01), ((0.9, 0.1, 0.2), 0.05), ((-0.2, 0.4, 0.5), 0.002)):
        rf = np.zeros((nu, nphi), dtype)
        K.cast_rays(rho, star, Q, dr, nsteps, dx, diag, L, rf)
        ref = ON.cast_rays(rho, [d(v) for v in star], Q, dr, nsteps, dx, diag, L, nu, nphi, sincos)
        np.testing.assert_array_equal(rf, ref)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_apply_front_and_smooth_match_numpy(okern32, okern64, dtype):
    K = okern32 if dtype == np.float32 else okern64
    d = np.dtype(dtype).type
    shape = (9, 8, 10)
    dx = d(1.0 / 10)
    rng = np.random.default_rng(8)
    rf = rng.uniform(0.05, 0.6, (8, 16)).astype(dtype)
    atan2 = (lambda y, x: K.lib.or_atan2f_det(float(y), float(x))) if dtype == np.float32 else math.atan2
    xi = rng.uniform(0, 0.5, shape).astype(dtype)
    df = rng.uniform(-0.2, 1.0, shape).astype(dtype)
    xi2, df2 = xi.copy(), df.copy()
    star = (0.45, 0.35, 0.55)
    K.apply_front(rf, star, dx, 1.2 * float(dx), xi, df)
    ON.apply_front(rf, star, dx, d(1.2 * float(dx)), xi2, df2, atan2)
    np.testing.assert_array_equal(xi, xi2)
    np.testing.assert_array_equal(df, df2)
    out = np.zeros(shape, dtype)
    K.smooth_xi(xi, out)
    np.testing.assert_array_equal(out, ON.smooth_xi(xi))


def test_max_speed_matches_numpy(okern64):
    U, xi = random_state((6, 7, 8), np.float64, 9)
    assert okern64.max_speed(*U, xi, 1 / 144, 1.0) == ON.max_speed(*U, xi, 1 / 144, 1.0)
