"""Pins of the pre-turbulence pieces: the stirring-field recipe (scenarios.stirring_field, the steps of
nebula3d.solenoidal) and the oracle's drive / add_torus / Mach sums (oracle/nebula_oracle.cpp), by closed forms."""
import math

import numpy as np
import pytest

F32 = np.float32


@pytest.fixture(scope="module")
def codec():
    from oracle.oracle import Codec
    return Codec()


def test_stirring_field_is_solenoidal_band_limited_unit_rms():
    from nebulaviewer_amd import scenarios as S
    shape = (20, 22, 26)
    v = S.stirring_field(shape, np.random.default_rng(4), 1.0, 4.0)
    assert math.isclose(math.sqrt(sum(float((c.astype(np.float64) ** 2).mean()) for c in v)), 1.0, rel_tol=1e-6)
    k = np.meshgrid(*[np.fft.fftfreq(n) * n for n in shape], indexing="ij")
    kk = np.sqrt(sum(q * q for q in k))
    V = [np.fft.fftn(c.astype(np.float64)) for c in v]
    power = sum(np.abs(x) ** 2 for x in V)
    out = (kk < 1.0) | (kk > 4.0)
    assert power[out].sum() < 1e-12 * power.sum()
    # divergence-free: k.V(k) = 0 for the projected modes, and the real part keeps it, since
    # k.conj(V(-k)) = -conj((-k).V(-k)) = 0; float32 storage leaves ~1e-14 of the power (an unprojected random
    # field has a third of it in k.V)
    div = sum(q * x for q, x in zip(k, V))
    assert (np.abs(div) ** 2).sum() < 1e-12 * (kk ** 2 * power).sum()


def test_drive_closed_forms(codec):
    """constant stirring fields: m += rho amp dt ((1 - blend) F + blend G) exactly as the float formula;
    affine fields are sampled exactly (trilinear); blend selects F (0) or G (1)"""
    n, dxs = 12, 1.0 / 6.0
    dx = 1.0 / n
    shape = (n, n, n)
    sshape = (8, 8, 8)
    rng = np.random.default_rng(5)
    rho = np.exp(rng.normal(0, 1, shape)).astype(F32)
    for blend, cF, cG in ((0.0, (1.0, -2.0, 0.5), (9.0, 9.0, 9.0)), (1.0, (9.0, 9.0, 9.0), (0.25, 1.0, -1.0)),
                          (0.3, (1.0, 0.0, 0.0), (0.0, 2.0, 0.0))):
        F = [np.full(sshape, F32(c)) for c in cF]
        G = [np.full(sshape, F32(c)) for c in cG]
        m = [np.zeros(shape, F32) for _ in range(3)]
        codec.drive(rho, m, F, G, blend, 7.0, 0.01, dx, dxs)
        for c in range(3):
            f = (1 - blend) * cF[c] + blend * cG[c]
            np.testing.assert_allclose(m[c], rho.astype(np.float64) * 7.0 * 0.01 * f, rtol=2e-6, atol=1e-12)
    # an affine field along x: F(x) = 2 + 3x, sampled at the cell centres
    xs = ((np.arange(sshape[0]) + 0.5) * dxs).astype(F32)
    Fx = np.broadcast_to((2 + 3 * xs)[:, None, None], sshape).astype(F32).copy()
    zero = np.zeros(sshape, F32)
    m = [np.zeros(shape, F32) for _ in range(3)]
    one = np.ones(shape, F32)
    codec.drive(one, m, [Fx, zero, zero], [zero, zero, zero], 0.0, 1.0, 1.0, dx, dxs)
    xc = (np.arange(n) + 0.5) * dx
    inside = (xc >= 0.5 * dxs) & (xc <= (sshape[0] - 0.5) * dxs)
    np.testing.assert_allclose(m[0][inside, 3, 3], 2 + 3 * xc[inside], rtol=1e-5)


def test_add_torus_and_mach(codec):
    shape = (6, 7, 8)
    rng = np.random.default_rng(6)
    rho = np.exp(rng.normal(0, 1, shape)).astype(F32)
    m = [rng.normal(0, 1, shape).astype(F32) for _ in range(3)]
    r0, m0 = rho.copy(), [c.copy() for c in m]
    codec.add_torus(rho, m, np.zeros(shape, F32))                 # nothing to add: bitwise unchanged
    assert np.array_equal(rho, r0) and all(np.array_equal(a, b) for a, b in zip(m, m0))
    tor = np.exp(rng.normal(0, 1, shape)).astype(F32)
    codec.add_torus(rho, m, tor)
    np.testing.assert_allclose(rho, r0.astype(np.float64) + tor, rtol=1e-6)
    # the momentum keeps the fraction of the new mass that was there before: m' = m r / (r + t)
    for a, b in zip(m, m0):
        np.testing.assert_allclose(a, b * r0 / (r0.astype(np.float64) + tor), rtol=3e-6, atol=1e-30)
    # Mach of a uniform flow: sum m^2/rho / sum rho = |v|^2
    rho_u = np.full(shape, 2.0, F32)
    v = (0.3, -0.4, 1.2)
    mu = [np.full(shape, F32(2.0 * c)) for c in v]
    a, b = codec.mach_sums(rho_u, mu)
    assert math.isclose(math.sqrt(a / b), math.sqrt(sum(c * c for c in v)), rel_tol=1e-6)
