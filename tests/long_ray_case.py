"""A long-ray case where the float recombination column stalls: dense gas brings the column close to Q, then
gas 2000x thinner adds per-sample increments near half an ulp of the column, so col + inc mostly rounds back
to col and a float ray overshoots the front by hundreds of cells, while an FP64 column finds it."""
import math

import numpy as np

SHAPE = (640, 24, 24)
X1 = 100              # dense region: x < X1 cells
N1, N2 = 30.0, 0.015  # densities inside and beyond
RAYS = (64, 256)      # nu, nphi: the rays nearest the +x axis stay inside the 24-cell cross-section
RAY = (32, 0)         # the probe ray (iu, ip): u = 1/64, phi = pi/256, within 1.1 degrees of +x


def geometry():
    nx, ny, nz = SHAPE
    dx = 1.0 / nx
    Lb = tuple(s * dx for s in SHAPE)
    diag = math.sqrt(sum(v * v for v in Lb)) + 0.6
    dr = 0.5 * dx
    nsteps = int(diag / dr) + 1
    rho = np.full(SHAPE, N2, np.float32)
    rho[:X1] = N1
    star = (0.5 * dx, 0.5 * ny * dx, 0.5 * nz * dx)
    return rho, star, dr, nsteps, dx, diag, Lb


def find_q(okern64, target_cells=2 * X1, iters=60):
    """Q for which the FP64 front of the probe ray lies target_cells from the star (bisection on Q: the
    front radius grows with Q)."""
    rho, star, dr, nsteps, dx, diag, Lb = geometry()
    rho64 = rho.astype(np.float64)
    rf = np.zeros(RAYS)

    def front(Q):
        okern64.cast_rays(rho64, star, Q, dr, nsteps, dx, diag, Lb, rf)
        return rf[RAY]

    lo = N1 * N1 * ((X1 - 2) * dx) ** 3 / 3      # front inside the dense region
    hi = N1 * N1 * ((X1 + 2) * dx) ** 3 / 3      # beyond anything the thin gas can add in the box
    for _ in range(iters):
        mid = 0.5 * (lo + hi)
        if front(mid) < target_cells * dx:
            lo = mid
        else:
            hi = mid
    return lo
