"""Closed-form checks that pin the CPU oracle's restatement of each reference kernel."""
import math

import numpy as np
import pytest

F32 = np.float32


# ------------------------------------------------------------------ limiters (nebula3d.py:40-46, sim.py:366-389)
@pytest.mark.parametrize("a,b,lim,expect", [
    (1, 2, 0, 1), (2, 1, 0, 1), (-1, -3, 0, -1), (1, -1, 0, 0), (0, 5, 0, 0), (2, 2, 0, 2), (-2, 2, 0, 0),
    (1, 3, 1, 1.5), (-1, -3, 1, -1.5), (1, -1, 1, 0), (0, 3, 1, 0),
    (1, 3, 2, 2), (1, 1.2, 2, 1.1), (-1, -3, 2, -2), (1, -2, 2, 0), (0.1, 10, 2, 0.2),
])
def test_limiters(okern64, a, b, lim, expect):
    assert okern64.fn("slope")(a, b, lim) == pytest.approx(expect, abs=1e-15)


def test_minmod_tie_returns_second(okern32):
    # |a| == |b|: the reference returns b (nebula3d.py:44-46); visible through the sign of a zero-free tie
    assert okern32.fn("slope")(3.0, 3.0, 0) == 3.0


# ------------------------------------------------------------------ fluxes (nebula3d.py:54-72, sim.py:397-420)
def flux(K, q4x4, c2b, c2c, axis, floor=2e-3, solver=1, lim=0):
    q = np.ascontiguousarray(np.asarray(q4x4, K.dtype).reshape(16))
    out = np.zeros(4, K.dtype)
    K.fn("face_flux")(q.ctypes.data, K.R(c2b), K.R(c2c), axis, K.R(floor), solver, lim, out.ctypes.data)
    return out


def physical_flux(rho, u, axis, c2):
    v = np.zeros(3)
    v[:] = u
    un = v[axis]
    F = np.array([rho * un, rho * v[0] * un, rho * v[1] * un, rho * v[2] * un])
    F[axis + 1] += rho * c2
    return F


@pytest.mark.parametrize("solver", [0, 1])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_flux_consistency(okern64, solver, axis):
    # uniform state: F(q,q,q,q) is the physical flux (HLL subsonic branch, Rusanov dissipation term vanishes)
    rho, u, c2 = 3.0, (0.3, -0.2, 0.1), 0.7
    q = [rho, *u] * 4
    F = flux(okern64, q, c2, c2, axis, solver=solver)
    np.testing.assert_allclose(F, physical_flux(rho, u, axis, c2), rtol=1e-14, atol=1e-15)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_hll_supersonic_upwind(okern32, axis):
    # un - c > 0 everywhere: SL >= 0, HLL returns the left flux exactly; un + c < 0: the right flux
    c2 = F32(0.25)
    for sgn, side in ((1, "L"), (-1, "R")):
        u = [0.0, 0.0, 0.0]
        u[axis] = sgn * 2.0
        qL = [2.0, *u]
        qR = [1.0, *u]
        q = qL + qL + qR + qR    # a = b = L, c = d = R: zero slopes, L = qb, R = qc
        F = flux(okern32, q, c2, c2, axis, solver=1)
        st = qL if side == "L" else qR
        expect = np.array(physical_flux(F32(st[0]), np.array(st[1:], F32), axis, c2), F32)
        np.testing.assert_array_equal(F, expect)


def test_rusanov_jump(okern64):
    # piecewise-constant jump (zero slopes): F = (FL+FR)/2 - amax/2 (UR-UL), amax = max(|uL|+cL, |uR|+cR)
    qL, qR, c2L, c2R = [2.0, 0.5, 0.1, 0.0], [1.0, -0.2, 0.0, 0.3], 0.04, 1.0
    F = flux(okern64, qL + qL + qR + qR, c2L, c2R, 0, solver=0)
    UL = np.array([2.0, 1.0, 0.2, 0.0])
    UR = np.array([1.0, -0.2, 0.0, 0.3])
    FL = physical_flux(2.0, qL[1:], 0, c2L)
    FR = physical_flux(1.0, qR[1:], 0, c2R)
    amax = max(0.5 + 0.2, 0.2 + 1.0)
    np.testing.assert_allclose(F, 0.5 * (FL + FR) - 0.5 * amax * (UR - UL), rtol=1e-14)


def test_hll_jump_formula(okern64):
    qL, qR, c2L, c2R = [2.0, 0.5, 0.1, 0.0], [1.0, -0.2, 0.0, 0.3], 0.04, 1.0
    F = flux(okern64, qL + qL + qR + qR, c2L, c2R, 0, solver=1)
    UL = np.array([2.0, 1.0, 0.2, 0.0])
    UR = np.array([1.0, -0.2, 0.0, 0.3])
    FL = physical_flux(2.0, qL[1:], 0, c2L)
    FR = physical_flux(1.0, qR[1:], 0, c2R)
    SL = min(0.5 - 0.2, -0.2 - 1.0)
    SR = max(0.5 + 0.2, -0.2 + 1.0)
    np.testing.assert_allclose(F, (SR * FL - SL * FR + SL * SR * (UR - UL)) / (SR - SL), rtol=1e-14)


def test_reconstruction_floor(okern64):
    # a steep density drop makes the reconstructed right state negative: it is floored at `floor`
    q = [10.0, 0, 0, 0, 10.0, 0, 0, 0, 0.01, 0, 0, 0, 0.01, 0, 0, 0]
    F = flux(okern64, q, 1.0, 1.0, 0, floor=2e-3, solver=1)
    assert np.isfinite(F).all()


# ------------------------------------------------------------------ sampler / slab test (nebula3d.py:195-244)
def test_sample_n_affine_and_outside(okern64):
    K = okern64
    n = (9, 7, 11)
    dx = 0.05
    i, j, k = np.meshgrid(*[np.arange(m, dtype=float) for m in n], indexing="ij")
    f = np.ascontiguousarray(1.5 + 0.3 * i - 0.7 * j + 0.2 * k)
    rng = np.random.default_rng(3)
    for _ in range(500):
        # interior of the cell-centre hull: trilinear interpolation of an affine field is exact
        g = np.array([rng.uniform(0, m - 1) for m in n])
        p = (g + 0.5) * dx
        v = K.fn("sample_n")(f.ctypes.data, *n, *p, dx)
        assert v == pytest.approx(1.5 + 0.3 * g[0] - 0.7 * g[1] + 0.2 * g[2], abs=1e-12)
    # outside [-1/2, n-1/2] in cell units: zero
    for p in ((-0.1 * dx, 0.1, 0.1), (0.1, 0.1, (n[2] + 0.1) * dx), (0.2, (n[1] + 0.01) * dx, 0.2)):
        assert K.fn("sample_n")(f.ctypes.data, *n, *p, dx) == 0.0
    # in the boundary band the clamped corners give the edge value
    c = np.ascontiguousarray(np.full(n, 4.25))
    assert K.fn("sample_n")(c.ctypes.data, *n, 0.01 * dx, 0.3, 0.2, dx) == 4.25


def test_box_range(okern64):
    K = okern64
    out = np.zeros(2)
    K.fn("box_range")(0.5, 0.5, 0.5, 1.0, 0.0, 0.0, 1.0, 2.0, 3.0, out.ctypes.data)
    assert out[0] == 0.0 and out[1] == pytest.approx(0.5)
    d = np.array([1.0, 2.0, 2.0]) / 3.0
    K.fn("box_range")(0.2, 0.2, 0.2, *d, 1.0, 1.0, 1.0, out.ctypes.data)
    assert out[0] == 0.0 and out[1] == pytest.approx(0.8 / (2 / 3))
    # from outside, towards the box: enters at the near face
    K.fn("box_range")(-1.0, 0.5, 0.5, 1.0, 0.0, 0.0, 1.0, 1.0, 1.0, out.ctypes.data)
    assert out[0] == pytest.approx(1.0) and out[1] == pytest.approx(2.0)
    # parallel to a slab it never crosses: a miss (t1 < t0)
    K.fn("box_range")(0.5, 2.0, 0.5, 1.0, 0.0, 0.0, 1.0, 1.0, 1.0, out.ctypes.data)
    assert out[1] < out[0]


# ------------------------------------------------------------------ rays (nebula3d.py:247-307)
@pytest.mark.parametrize("dens,Q", [(1.0, 0.15 ** 3 / 3), (2.0, 0.05), (0.5, 0.002)])
def test_stromgren_uniform(okern32, dens, Q):
    K = okern32
    n = 64
    dx = 1.0 / n
    rho = np.full((n, n, n), dens, F32)
    rf = np.zeros((32, 64), F32)
    diag = math.sqrt(3) + 0.6
    dr = 0.5 * dx
    K.cast_rays(rho, (0.5, 0.5, 0.5), Q, dr, int(diag / dr) + 1, dx, diag, (1, 1, 1), rf)
    R = (3 * Q / dens ** 2) ** (1 / 3)
    if R < 0.45:
        # midpoint rule + linear last-step interpolation: O(dr^2/R^2) relative error
        np.testing.assert_allclose(rf, R, rtol=2 * (dr / R) ** 2 + 1e-6)
    else:
        # the rays leave the box before the column reaches Q: rf stays at rcap = diag
        assert (rf[rf > 0.9] == np.float32(diag)).any()


def test_stromgren_leaves_box(okern32):
    K = okern32
    n = 32
    dx = 1.0 / n
    rho = np.full((n, n, n), 0.01, F32)
    rf = np.zeros((8, 16), F32)
    diag = math.sqrt(3) + 0.6
    K.cast_rays(rho, (0.5, 0.5, 0.5), 1.0, 0.5 * dx, int(diag / (0.5 * dx)) + 1, dx, diag, (1, 1, 1), rf)
    assert (rf == np.float32(diag)).all()


def test_apply_front_constant_radius(okern64):
    K = okern64
    n = (12, 10, 14)
    dx = 1.0 / 12
    R0, w, star = 0.31, 1.2 * dx, (0.41, 0.47, 0.52)
    rf = np.full((16, 32), R0)
    xi = np.zeros(n)
    df = np.full(n, 1e3)
    K.apply_front(rf, star, dx, w, xi, df)
    i, j, k = np.meshgrid(*[np.arange(m) for m in n], indexing="ij")
    r = np.sqrt(((i + 0.5) * dx - star[0]) ** 2 + ((j + 0.5) * dx - star[1]) ** 2 + ((k + 0.5) * dx - star[2]) ** 2)
    np.testing.assert_allclose(xi, np.clip((R0 - r) / w + 0.5, 0, 1), atol=1e-13)
    np.testing.assert_allclose(df, r - R0, atol=1e-13)
    # a second star can only raise xi and lower dfront
    xi2, df2 = xi.copy(), df.copy()
    K.apply_front(np.full((16, 32), 0.1), (0.1, 0.1, 0.1), dx, w, xi2, df2)
    assert (xi2 >= xi).all() and (df2 <= df).all()


def test_smooth_xi_brute_force(okern64):
    K = okern64
    n = (5, 6, 7)
    src = np.random.default_rng(4).uniform(size=n)
    dst = np.zeros(n)
    K.smooth_xi(src, dst)
    cl = lambda v, m: min(max(v, 0), m - 1)
    for i in range(n[0]):
        for j in range(n[1]):
            for k in range(n[2]):
                nb = [src[cl(i - 1, n[0]), j, k], src[cl(i + 1, n[0]), j, k], src[i, cl(j - 1, n[1]), k],
                      src[i, cl(j + 1, n[1]), k], src[i, j, cl(k - 1, n[2])], src[i, j, cl(k + 1, n[2])]]
                assert dst[i, j, k] == pytest.approx(0.5 * src[i, j, k] + 0.5 * sum(nb) / 6, abs=1e-15)


# ------------------------------------------------------------------ FUV (nebula3d.py:320-388)
def test_fuv_uniform_attenuation(okern64):
    K = okern64
    n = 32
    dx = 1.0 / n
    dens, kappa = 2.0, 3.0
    rho = np.full((n, n, n), dens)
    xi = np.zeros((n, n, n))
    diag = math.sqrt(3) + 0.6
    dr = 0.5 * dx
    nr = 24
    binr = diag / nr
    att = np.zeros((nr, 8, 16))
    star = (0.5, 0.5, 0.5)
    K.cast_fuv(rho, xi, star, kappa, dr, int(diag / dr) + 1, dx, binr, (1, 1, 1), att)
    # inside the box, bin b holds exp(-tau) at the first sample with r >= (b+1) binr: tau = kappa n (m+1) dr
    for b in range(4):
        m = math.ceil((b + 1) * binr / dr - 0.5)
        np.testing.assert_allclose(att[b], math.exp(-kappa * dens * (m + 1) * dr), rtol=1e-12)
    # kappa = 0: no attenuation; g0 = lum / max(r^2, 4e-4) exactly up to rounding of the weights
    att1 = np.ones((nr, 8, 16))
    g0 = np.zeros((n, n, n))
    K.apply_fuv(att1, star, 1.7, dx, binr, g0)
    i, j, k = np.meshgrid(*[np.arange(n)] * 3, indexing="ij")
    r2 = ((i + 0.5) * dx - 0.5) ** 2 + ((j + 0.5) * dx - 0.5) ** 2 + ((k + 0.5) * dx - 0.5) ** 2
    np.testing.assert_allclose(g0, 1.7 / np.maximum(r2, 4e-4), rtol=1e-13)


# ------------------------------------------------------------------ fp16 store (sim.py:511-514), CFL (154-162)
def test_to_half_matches_numpy(okern32):
    rng = np.random.default_rng(5)
    parts = [rng.standard_normal(100000) * 10.0 ** rng.uniform(-9, 5.5, 100000),
             np.float32([0.0, -0.0, 65504, 65519.99, 65520, 70000, -70000, 2 ** -24, 2 ** -25, 3 * 2 ** -26,
                         2 ** -25 * 1.0000001, 2 ** -14, 6.1e-5, 1e30, -1e-30, np.inf, -np.inf])]
    # exact ties between adjacent halves (round to even)
    h = np.arange(0, 0x7BFF, 7, dtype=np.uint16).view(np.float16).astype(np.float32)
    h2 = np.arange(1, 0x7C00, 7, dtype=np.uint16).view(np.float16).astype(np.float32)
    parts.append((h[: len(h2)] + h2[: len(h)]) * np.float32(0.5))
    x = np.concatenate(parts).astype(F32)
    got = okern32.to_half(x)
    np.testing.assert_array_equal(got.view(np.uint16), x.astype(np.float16).view(np.uint16))


def test_max_speed_brute_force(okern32):
    rng = np.random.default_rng(6)
    n = (7, 9, 5)
    rho = rng.uniform(0.01, 3, n).astype(F32)
    m = [rng.standard_normal(n).astype(F32) for _ in range(3)]
    xi = rng.uniform(0, 1, n).astype(F32)
    cn2, ci2 = F32((1 / 12) ** 2), F32(1.0)
    got = okern32.max_speed(rho, *m, xi, cn2, ci2)
    ref = (((np.abs(m[0] / rho) + np.abs(m[1] / rho)) + np.abs(m[2] / rho)) + np.sqrt(cn2 + xi * (ci2 - cn2))).max()
    assert np.float32(got) == ref
