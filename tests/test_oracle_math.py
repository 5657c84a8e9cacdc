"""The deterministic float transcendentals (nv_math.h) shared by the HIP kernels and the CPU oracle, bounded
against the platform libm in double precision (math module). These are what the reference's wp.sin, wp.cos,
wp.atan2 and wp.exp calls (nebula3d.py:256-257, 287, 345) compute."""
import math

import numpy as np
import pytest


def ulp_dist(a, b):
    a = np.asarray(a, np.float32).view(np.int32).astype(np.int64)
    b = np.asarray(b, np.float32).view(np.int32).astype(np.int64)
    a = np.where(a < 0, -(a & 0x7FFFFFFF), a)
    b = np.where(b < 0, -(b & 0x7FFFFFFF), b)
    return np.abs(a - b)


@pytest.fixture(scope="module")
def lib():
    from oracle.oracle import load
    return load()


def test_sincos_ray_angles(lib):
    # every angle a ray grid uses, plus random angles in [0, 2pi)
    xs = []
    for nphi in (256, 1024, 2048, 320, 512):
        ip = np.arange(nphi, dtype=np.float32)
        xs.append((ip + np.float32(0.5)) * np.float32(2) * np.float32(math.pi) / np.float32(nphi))
    xs.append(np.random.default_rng(0).uniform(0, 2 * math.pi, 20000).astype(np.float32))
    x = np.concatenate(xs)
    s = np.array([lib.or_sinf_det(float(v)) for v in x], np.float32)
    c = np.array([lib.or_cosf_det(float(v)) for v in x], np.float32)
    s_ref = np.array([math.sin(float(v)) for v in x])
    c_ref = np.array([math.cos(float(v)) for v in x])
    # at most one ulp from the correctly rounded value (the double evaluation is ~1e-16 accurate)
    assert ulp_dist(s, s_ref.astype(np.float32)).max() <= 1
    assert ulp_dist(c, c_ref.astype(np.float32)).max() <= 1
    # and absolutely within float resolution near the zeros of sin/cos
    assert np.abs(s - s_ref).max() < 1e-7
    assert np.abs(c - c_ref).max() < 1e-7


def test_atan2_quadrants_and_zeros(lib):
    # float-only evaluation: within 3 ulp of the correctly rounded atan2 on random pairs of every scale and sign,
    # on the octant diagonal, around the reduction's switch at tan(pi/12) and its cancellation at 1/sqrt3, and
    # on the cell-centre offsets the ionize pass feeds it
    rng = np.random.default_rng(1)
    n = 400_000
    sgn = lambda: rng.choice([-1.0, 1.0], n)
    mag = rng.uniform(1e-3, 10.0, n)
    cases = [
        (rng.standard_normal(n) * 10.0 ** rng.integers(-6, 6, n), rng.standard_normal(n) * 10.0 ** rng.integers(-6, 6, n)),
        (mag * (1 + rng.uniform(-1e-3, 1e-3, n)) * sgn(), mag * sgn()),
        (mag * 0.2679491924311227 * (1 + rng.uniform(-1e-4, 1e-4, n)) * sgn(), mag * sgn()),
        (mag * 0.5773502691896258 * (1 + rng.uniform(-1e-3, 1e-3, n)) * sgn(), mag * sgn()),
    ]
    dx = np.float32(1.0 / 512)
    ij = rng.integers(0, 512, (2, n)).astype(np.float32)
    cases.append(((ij[1] + np.float32(0.5)) * dx - np.float32(0.37), (ij[0] + np.float32(0.5)) * dx - np.float32(0.5)))
    worst = 0
    for y, x in cases:
        y = np.ascontiguousarray(y, np.float32)
        x = np.ascontiguousarray(x, np.float32)
        got = np.empty_like(y)
        lib.or_atan2f_det_n(y.ctypes.data, x.ctypes.data, got.ctypes.data, n, 4)
        ref = np.arctan2(y.astype(np.float64), x.astype(np.float64)).astype(np.float32)
        worst = max(worst, int(ulp_dist(got, ref).max()))
    assert worst <= 3
    # the scalar entry point is the same function
    y, x = np.float32([0.5, -3.0, 1e-3]), np.float32([3.0, 0.25, -7.0])
    many = np.empty_like(y)
    lib.or_atan2f_det_n(y.ctypes.data, x.ctypes.data, many.ctypes.data, 3, 1)
    assert [np.float32(lib.or_atan2f_det(float(a), float(b))) for a, b in zip(y, x)] == list(many)
    # axes and C99 signed-zero conventions
    assert lib.or_atan2f_det(0.0, 1.0) == 0.0 and math.copysign(1, lib.or_atan2f_det(-0.0, 1.0)) < 0
    assert lib.or_atan2f_det(0.0, -1.0) == np.float32(math.pi)
    assert lib.or_atan2f_det(-0.0, -1.0) == -np.float32(math.pi)
    assert lib.or_atan2f_det(0.0, -0.0) == np.float32(math.pi)
    assert lib.or_atan2f_det(1.0, 0.0) == np.float32(math.pi / 2)
    assert lib.or_atan2f_det(-2.0, 0.0) == -np.float32(math.pi / 2)
    assert ulp_dist(lib.or_atan2f_det(1.0, 1.0), np.float32(math.pi / 4)) <= 1


def test_exp_range(lib):
    x = np.concatenate([-np.random.default_rng(2).uniform(0, 110, 20000), np.linspace(-1, 0, 1001)]).astype(np.float32)
    got = np.array([lib.or_expf_det(float(v)) for v in x], np.float32)
    ref = np.array([math.exp(float(v)) for v in x]).astype(np.float32)
    normal = ref >= np.float32(2.0 ** -126)
    assert ulp_dist(got[normal], ref[normal]).max() <= 1
    # float subnormal results: within one subnormal step; below the range: +0
    assert np.abs(got[~normal] - ref[~normal]).max() <= 2.0 ** -149
    assert lib.or_expf_det(-200.0) == 0.0 and lib.or_expf_det(0.0) == 1.0
