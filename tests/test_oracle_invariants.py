"""Invariants and exact solutions that pin the oracle's hydro stage (nebula3d.py:76-151, sim.py:423-498)
beyond the closed-form flux checks: steady uniform states, mirror symmetry, conservation telescoping in FP64,
and convergence to the exact isothermal Riemann solution."""
import math

import numpy as np
import pytest

import oracle_numpy as ON


def stage(K, U, xi, dt, dx, solver=1, lim=0, a=0.0, b=1.0, U0=None, cn2=1 / 144, floor=2e-3):
    d = K.dtype.type
    shape = U[0].shape
    out = [np.zeros(shape, K.dtype) for _ in range(4)]
    z = np.zeros((1, 4), K.dtype)
    K.hydro_stage(U, xi, U if U0 is None else U0, out, shape, z, 0, d(0), z, d(0.9), d(0.04), d(1), d(dt), d(dx),
                  d(a), d(b), d(floor), d(cn2), d(1.0), d(6.0), solver, lim)
    return out


@pytest.mark.parametrize("solver,lim", [(1, 0), (1, 1), (1, 2), (0, 0)])
def test_uniform_flow_is_steady(okern32, solver, lim):
    # every face carries the same flux, so dU = 0 exactly; clamped (outflow) ghosts keep it so at the walls
    shape = (7, 9, 11)
    U = [np.full(shape, v, np.float32) for v in (3.0, 0.6, -0.3, 0.15)]
    xi = np.full(shape, 0.25, np.float32)
    out = stage(okern32, U, xi, 1e-3, 1 / 11, solver, lim)
    for c in range(4):
        np.testing.assert_array_equal(out[c], U[c])


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("solver", [0, 1])
def test_mirror_symmetry(okern64, axis, solver):
    # a state mirrored along one axis (with the normal momentum negated) evolves into the mirrored state
    rng = np.random.default_rng(30 + axis)
    shape = (10, 12, 14)
    rho = rng.lognormal(0, 0.5, shape)
    m = [rho * rng.normal(0, 0.3, shape) for _ in range(3)]
    xi = rng.uniform(0, 1, shape)
    flip = lambda f: np.ascontiguousarray(np.flip(f, axis))
    U = [rho, *m]
    Uf = [flip(rho)] + [(-flip(m[c]) if c == axis else flip(m[c])) for c in range(3)]
    o = stage(okern64, U, xi, 2e-3, 1 / 14, solver)
    of = stage(okern64, Uf, flip(xi), 2e-3, 1 / 14, solver)
    np.testing.assert_allclose(flip(of[0]), o[0], rtol=1e-13, atol=1e-15)
    for c in range(3):
        sgn = -1 if c == axis else 1
        np.testing.assert_allclose(sgn * flip(of[c + 1]), o[c + 1], rtol=1e-12, atol=1e-13)


def test_mass_telescoping_fp64(okern64):
    # sum over cells of the mass update = -dt/dx (outflow through the 6 boundary planes); the interior face
    # fluxes cancel pairwise because each face is evaluated from identical inputs on both sides
    rng = np.random.default_rng(31)
    shape = (9, 10, 12)
    dx, dt = 1 / 12, 1e-3
    rho = rng.lognormal(0, 0.5, shape) + 0.1
    m = [rho * rng.normal(0, 0.4, shape) for _ in range(3)]
    xi = rng.uniform(0, 1, shape)
    cn2 = 1 / 144
    out = stage(okern64, [rho, *m], xi, dt, dx, cn2=cn2)
    r = np.maximum(rho, 2e-3)
    q = [r, m[0] / r, m[1] / r, m[2] / r]
    c2 = cn2 + xi * (1 - cn2)
    net = 0.0
    for axis in range(3):
        def sl(i):
            s = [slice(None)] * 3
            s[axis] = i
            return tuple(s)
        hi = [[qq[sl(-2)] for qq in q], [qq[sl(-1)] for qq in q]]
        lo = [[qq[sl(0)] for qq in q], [qq[sl(1)] for qq in q]]
        Fh = ON._flux(hi[0], hi[1], hi[1], hi[1], c2[sl(-1)], c2[sl(-1)], axis, 2e-3, 1, 0)
        Fl = ON._flux(lo[0], lo[0], lo[0], lo[1], c2[sl(0)], c2[sl(0)], axis, 2e-3, 1, 0)
        net += Fh[0].sum() - Fl[0].sum()
    dM = out[0].sum() - rho.sum()
    assert abs(dM - (-dt / dx * net)) < 1e-12 * rho.sum()


# ------------------------------------------------------------------ exact isothermal Riemann problem
def exact_isothermal(rl, ul, rr, ur, c, xs):
    """Exact solution at similarity coordinates xs = x/t: shocks (Rankine-Hugoniot) or rarefactions."""
    def f(r, rk):
        return c * (r - rk) / math.sqrt(r * rk) if r > rk else c * math.log(r / rk)

    def df(r, rk):
        return c * (r + rk) / (2 * r * math.sqrt(r * rk)) if r > rk else c / r
    rs = math.sqrt(rl * rr)
    for _ in range(100):
        g = f(rs, rl) + f(rs, rr) + (ur - ul)
        rs = max(rs - g / (df(rs, rl) + df(rs, rr)), 1e-12)
        if abs(g) < 1e-14:
            break
    us = 0.5 * (ul + ur) + 0.5 * (f(rs, rr) - f(rs, rl))
    rho = np.empty_like(xs)
    u = np.empty_like(xs)
    for n, s in enumerate(xs):
        if s < us:
            if rs > rl:   # left shock
                S = ul - c * math.sqrt(rs / rl)
                rho[n], u[n] = (rl, ul) if s < S else (rs, us)
            else:         # left rarefaction
                if s < ul - c:
                    rho[n], u[n] = rl, ul
                elif s > us - c:
                    rho[n], u[n] = rs, us
                else:
                    u[n] = s + c
                    rho[n] = rl * math.exp((ul - u[n]) / c)
        else:
            if rs > rr:   # right shock
                S = ur + c * math.sqrt(rs / rr)
                rho[n], u[n] = (rr, ur) if s > S else (rs, us)
            else:
                if s > ur + c:
                    rho[n], u[n] = rr, ur
                elif s < us + c:
                    rho[n], u[n] = rs, us
                else:
                    u[n] = s - c
                    rho[n] = rr * math.exp((u[n] - ur) / c)
    return rho, u


def run_riemann(K, n, rl, ul, rr, ur, t_end, solver=1):
    dx = 1.0 / n
    x = (np.arange(n) + 0.5) * dx
    rho = np.where(x < 0.5, rl, rr).reshape(n, 1, 1)
    mx = np.where(x < 0.5, rl * ul, rr * ur).reshape(n, 1, 1)
    U = [np.ascontiguousarray(rho), np.ascontiguousarray(mx), np.zeros((n, 1, 1)), np.zeros((n, 1, 1))]
    xi = np.ones((n, 1, 1))           # fully ionized: c^2 = ci2 = 1
    t = 0.0
    while t < t_end - 1e-14:
        smax = K.max_speed(*U, xi, 1 / 144, 1.0)
        dt = min(0.3 * dx / smax, t_end - t)
        U1 = stage(K, U, xi, dt, dx, solver)
        U = stage(K, U1, xi, dt, dx, solver, a=0.5, b=0.5, U0=U)
        t += dt
    return x, U[0][:, 0, 0], U[1][:, 0, 0] / U[0][:, 0, 0]


@pytest.mark.parametrize("case", [(4.0, 0.0, 1.0, 0.0), (1.0, 1.5, 1.0, -1.5), (1.0, -1.0, 1.0, 1.0),
                                  (3.0, 0.5, 0.5, -0.2)])
@pytest.mark.parametrize("solver", [1, 0])
def test_riemann_converges_to_exact(okern64, case, solver):
    rl, ul, rr, ur = case
    t_end = 0.15
    errs = []
    for n in (200, 400):
        x, r, u = run_riemann(okern64, n, rl, ul, rr, ur, t_end, solver)
        re, _ = exact_isothermal(rl, ul, rr, ur, 1.0, (x - 0.5) / t_end)
        errs.append(np.abs(r - re).mean() / np.abs(re).mean())
    # first order at discontinuities: the L1 error falls with resolution and is small at 400 cells
    assert errs[1] < 0.8 * errs[0], errs
    assert errs[1] < 0.02, errs
