"""Pins of the oracle's restatement of the consumers' kernels (oracle/codec_oracle.cpp): the timeline codec
(timeline.py:19-89), the shock proxy (nebula3d.py:182-191), the render pack and prefilter (render.py:121-130,
176-204, 281-298). Each check is a property or a closed form, not the kernel's formula retyped."""
import numpy as np
import pytest

F32 = np.float32


@pytest.fixture(scope="module")
def codec():
    from oracle.oracle import Codec
    return Codec()


def fields(shape, seed):
    rng = np.random.default_rng(seed)
    rho = np.exp(rng.normal(0, 2, shape)).astype(F32)
    m = [(rho * rng.normal(0, 2, shape)).astype(F32) for _ in range(3)]
    xi = rng.uniform(-0.2, 1.2, shape).astype(F32)
    df = rng.normal(0, 3, shape).astype(F32)
    g0 = np.exp(rng.normal(0, 3, shape)).astype(F32)
    return rho, m, xi, df, g0


def test_timeline_pack_is_saturated_binary16(codec):
    """every stored field is numpy's round-to-nearest-even binary16 of the value saturated at 65000 (rho, g0
    above; m, dfront both ways), xi is int(255 clamp(xi, 0, 1) + 1/2)"""
    shape = (7, 9, 11)
    rho, m, xi, df, g0 = fields(shape, 1)
    rho.flat[:5] = [7e4, 65000, 64991, 1e6, 0.0]
    m[0].flat[:4] = [-7e4, 7e4, -65000.5, 3.0]
    g0.flat[:2] = [1e9, 65520]
    out = codec.timeline_pack(rho, *m, xi, df, g0)
    h = lambda a: a.astype(np.float16).view(np.uint16)
    np.testing.assert_array_equal(out[0], h(np.minimum(rho, F32(65000))))
    for c in range(3):
        np.testing.assert_array_equal(out[1 + c], h(np.clip(m[c], F32(-65000), F32(65000))))
    np.testing.assert_array_equal(out[4], (np.clip(xi, 0, 1) * F32(255) + F32(0.5)).astype(np.int32).astype(np.uint8))
    np.testing.assert_array_equal(out[5], h(np.clip(df, F32(-65000), F32(65000))))
    np.testing.assert_array_equal(out[6], h(np.minimum(g0, F32(65000))))
    assert np.float16(65000) == np.float16(64992)        # the saturated code itself
    assert (out[0].view(np.float16).astype(F32) <= 65000).all()


def test_timeline_round_trip(codec):
    """values binary16 and 8 bits represent come back exactly; rho below the floor comes back as the floor"""
    shape = (5, 6, 7)
    rng = np.random.default_rng(2)
    h = lambda: rng.normal(0, 50, shape).astype(np.float16).astype(F32)
    rho = (np.abs(h()) + F32(1)).astype(np.float16).astype(F32)
    rho.flat[0] = F32(1e-4)
    m = [h(), h(), h()]
    xi = (rng.integers(0, 256, shape) / F32(255)).astype(F32)
    df, g0 = h(), np.abs(h())
    st = codec.timeline_pack(rho, *m, xi, df, g0)
    back = codec.timeline_unpack(st, 2e-3)
    expect_rho = np.maximum(rho, F32(2e-3))
    for got, want in zip(back, [expect_rho] + m + [xi, df, g0]):
        np.testing.assert_array_equal(got, want)
    assert back[0].flat[0] == F32(2e-3)


def test_shock_closed_forms(codec):
    """zero for uniform and expanding flows; -div(v) rho on a linear compression (central differences are exact
    on linear fields; one-sided at the clamped faces, where the difference spans one cell instead of two)"""
    n = 12
    dx = 1.0 / n
    shape = (n, n, n)
    x = ((np.arange(n) + 0.5) * dx).astype(F32)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    rho = np.full(shape, 2.0, F32)
    zero = np.zeros(shape, F32)
    assert not codec.shock_field(rho, zero, zero, zero, dx).any()
    a, b, c = 3.0, 1.5, 0.5
    expand = [(rho * F32(s) * q).astype(F32) for s, q in ((a, X), (b, Y), (c, Z))]
    assert not codec.shock_field(rho, *expand, dx).any()
    comp = [(-e).astype(F32) for e in expand]
    out = codec.shock_field(rho, *comp, dx)
    inner = out[1:-1, 1:-1, 1:-1]
    np.testing.assert_allclose(inner, (a + b + c) * 2.0, rtol=2e-5)
    # a corner cell: every axis one-sided (half the difference over the same 2 dx)
    np.testing.assert_allclose(out[0, 0, 0], (a + b + c) * 2.0 / 2, rtol=2e-5)


def pp_vec(c_ha=(0, 0, 0), c_o3=(0, 0, 0), c_s2=(0, 0, 0), c_dg=(0, 0, 0), c_sh=(0, 0, 0), gs=10.0, ambient=0.0,
           edge=0.0, inv_L=(1, 1, 1), dx=0.1):
    return np.array(list(c_ha) + list(c_o3) + list(c_s2) + list(c_dg) + list(c_sh) + [gs, ambient, edge] +
                    list(inv_L) + [dx], F32)


def unpack16(out):
    return out.view(np.float16).astype(np.float64)


def test_pack_volume_single_terms(codec):
    """each emission term on its own, against its closed form (to binary16 rounding)"""
    shape = (3, 3, 3)
    rho = np.full(shape, 2.0, F32)
    zero = np.zeros(shape, F32)
    xi = np.full(shape, 0.5, F32)
    df = np.full(shape, -1.0, F32)      # deep inside the front: both [O III] and [S II] weights fully on / off
    g0 = np.full(shape, 30.0, F32)
    # H-alpha: (rho xi)^2 c_ha
    out = unpack16(codec.pack_volume(rho, zero, zero, zero, xi, df, g0, pp_vec(c_ha=(1, 2, 0.5)), 0))
    np.testing.assert_allclose(out[1, 1, 1, :3], np.array([1, 2, 0.5]) * 1.0, rtol=1e-3)
    # dust = rho (1 - xi) + 0.08 rho xi = 1.08; 4th channel
    np.testing.assert_allclose(out[1, 1, 1, 3], 1.08, rtol=1e-3)
    # [O III] weight sstep(-0.03, -0.12, df): 1 for df <= -0.12, 0 for df >= -0.03, 1/2 at the midpoint
    for d, w in ((-1.0, 1.0), (0.0, 0.0), (-0.075, 0.5)):
        out = unpack16(codec.pack_volume(rho, zero, zero, zero, xi, np.full(shape, d, F32), g0,
                                         pp_vec(c_o3=(1, 0, 0)), 0))
        np.testing.assert_allclose(out[1, 1, 1, 0], w * 1.0, rtol=1e-3, atol=1e-6)
    # [S II] weight sstep(-0.05, -0.005, df): 0 below -0.05, 1 above -0.005
    for d, w in ((-1.0, 0.0), (0.5, 1.0)):
        out = unpack16(codec.pack_volume(rho, zero, zero, zero, xi, np.full(shape, d, F32), g0,
                                         pp_vec(c_s2=(0, 1, 0)), 0))
        np.testing.assert_allclose(out[1, 1, 1, 1], w * 1.0, rtol=1e-3, atol=1e-6)
    # dust x G0 with saturation g / (1 + g / gs) + ambient
    out = unpack16(codec.pack_volume(rho, zero, zero, zero, xi, df, g0, pp_vec(c_dg=(0, 0, 1), gs=10, ambient=0.25),
                                     0))
    np.testing.assert_allclose(out[1, 1, 1, 2], 1.08 * (30 / (1 + 3) + 0.25), rtol=1e-3)
    # saturation of the stored value at 60000
    big = np.full(shape, 1e4, F32)
    out = unpack16(codec.pack_volume(big, zero, zero, zero, np.ones(shape, F32), df, g0, pp_vec(c_ha=(1, 0, 0)), 0))
    assert out[1, 1, 1, 0] == float(np.float16(60000))


def test_pack_volume_edge_fade(codec):
    """the face fade sstep(0, e, u) sstep(0, e, 1 - u) per axis: 1 in the interior, 0 at u <= 0, symmetric"""
    n = 20
    shape = (n, n, n)
    rho = np.ones(shape, F32)
    zero = np.zeros(shape, F32)
    xi = np.ones(shape, F32)
    pp = pp_vec(c_ha=(1, 0, 0), edge=0.2, dx=1.0 / n, inv_L=(1, 1, 1))
    out = unpack16(codec.pack_volume(rho, zero, zero, zero, xi, zero - 1, zero, pp, 0))[..., 0]
    assert out[n // 2, n // 2, n // 2] == 1.0                       # u in [e, 1 - e] on every axis
    u0 = 0.5 / n                                                    # first cell centre
    t = u0 / 0.2
    s = t * t * (3 - 2 * t)
    np.testing.assert_allclose(out[0, n // 2, n // 2], s, rtol=2e-3)
    np.testing.assert_allclose(out[0, 0, 0], s ** 3, rtol=5e-3)
    np.testing.assert_allclose(out, out[::-1, :, :], rtol=1e-3)      # symmetric (u and 1 - u rounded apart)


def test_smooth_axis(codec):
    """constants stay; a delta becomes [1 4 1]/6 along the axis only; the ends are clamped"""
    shape = (5, 6, 7)
    const = np.full(shape + (4,), np.float16(3.25)).view(np.uint16)
    for axis in range(3):
        np.testing.assert_array_equal(codec.smooth_axis(const, axis), const)
    delta = np.zeros(shape + (4,), np.float16)
    delta[2, 3, 3, 0] = np.float16(6.0)
    d16 = delta.view(np.uint16)
    for axis in range(3):
        out = codec.smooth_axis(d16, axis).view(np.float16).astype(np.float64)
        assert out[2, 3, 3, 0] == 4.0
        idx = [2, 3, 3]
        idx[axis] -= 1
        assert out[tuple(idx) + (0,)] == 1.0
        idx[axis] += 2
        assert out[tuple(idx) + (0,)] == 1.0
        assert out.sum() == 6.0                     # only channel 0 is set: 4 + 1 + 1
        assert out[..., 1:].sum() == 0.0
    edge = np.zeros(shape + (4,), np.float16)
    edge[0, 0, 0, 2] = np.float16(6.0)
    out = codec.smooth_axis(edge.view(np.uint16), 0).view(np.float16).astype(np.float64)
    assert out[0, 0, 0, 2] == 5.0 and out[1, 0, 0, 2] == 1.0   # clamped: the missing neighbour is the cell itself
