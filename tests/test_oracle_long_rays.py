"""The FP64 long-ray column of the oracle (nv_cast_rays_wide's checker) pinned on cases with known answers.

  - rays shorter than 512 cells are bit-identical to the float cast_rays (nebula3d.py:247-270);
  - on a stalled-column case (tests/long_ray_case.py) the float ray overshoots the front by more than 50 cells,
    while the FP64 column lands within 12 cells of the all-FP64 oracle;
  - on a uniform cube the long diagonal rays agree with the closed form (3Q/n^2)^(1/3) to O(dr^2).
"""
import math

import numpy as np

import long_ray_case as LR


def test_short_rays_identical(okern32):
    shape = (40, 36, 44)
    rng = np.random.default_rng(3)
    rho = rng.lognormal(0, 1.0, shape).astype(np.float32)
    dx = 1.0 / 44
    Lb = tuple(n * dx for n in shape)
    diag = math.sqrt(sum(v * v for v in Lb)) + 0.6
    a, b = np.zeros((32, 64), np.float32), np.zeros((32, 64), np.float32)
    for Q in (0.003, 0.05, 10.0):
        okern32.cast_rays(rho, (0.4, 0.35, 0.5), Q, 0.5 * dx, int(diag / (0.5 * dx)) + 1, dx, diag, Lb, a)
        okern32.cast_rays(rho, (0.4, 0.35, 0.5), Q, 0.5 * dx, int(diag / (0.5 * dx)) + 1, dx, diag, Lb, b, wide=True)
        np.testing.assert_array_equal(a, b)


def test_stalled_column_found_in_fp64(okern32, okern64):
    Q = LR.find_q(okern64)
    rho, star, dr, nsteps, dx, diag, Lb = LR.geometry()
    out = {}
    for name, K, wide in (("f32", okern32, False), ("wide", okern32, True), ("f64", okern64, False)):
        rf = np.zeros(LR.RAYS, K.dtype)
        K.cast_rays(rho.astype(K.dtype), star, Q, dr, nsteps, dx, diag, Lb, rf, wide=wide)
        out[name] = float(rf[LR.RAY])
    # the FP64 front is where the bisection put it, in the thin gas
    assert abs(out["f64"] - 2 * LR.X1 * dx) < 0.01 * dx
    # the float column all but stalls past the dense region (increments round away) and overshoots
    assert out["f32"] - out["f64"] > 50 * dx, out
    # the FP64 column of the float sampler finds the same front up to its float sampling: near the stall a
    # relative column change of 3e-7 (the float positions of the density step) moves the front ~10 cells
    assert abs(out["wide"] - out["f64"]) < 12 * dx, out


def test_uniform_cube_diagonal(okern32):
    n = 320
    shape = (n, n, n)
    dx = 1.0 / n
    Lb = tuple(s * dx for s in shape)
    diag = math.sqrt(sum(v * v for v in Lb)) + 0.6
    dr = 0.5 * dx
    rho = np.full(shape, 1.0, np.float32)
    rf_true = 1.55                      # 496 cells: only near-diagonal rays from the corner reach it
    Q = rf_true ** 3 / 3
    rf = np.zeros((64, 128), np.float32)
    okern32.cast_rays(rho, (0.5 * dx,) * 3, Q, dr, int(diag / dr) + 1, dx, diag, Lb, rf, wide=True)
    found = rf < diag - 1e-3
    assert found.sum() >= 4
    # the star sits at the centre of the corner cell, so the column starts at r = 0; the midpoint rule of
    # r^2 undershoots by dr^2 r / 12 per unit length and the front moves out by about dr^2 / (12 rf)
    assert np.max(np.abs(rf[found] - rf_true)) < 1e-3 * dx
