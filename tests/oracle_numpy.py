"""Independent NumPy restatement of the reference step kernels, used only to cross-check the C++ oracle.

Written separately from oracle/nebula_oracle.cpp (vectorised over the grid with edge padding for the clamped
ghosts of nebula3d.ldc, instead of per-cell loops with clamped indices), following the same reference lines:
hydro_stage / hydro_stage_hll (nebula3d.py:76-151, sim.py:423-498), face_flux (nebula3d.py:54-72), flux_hll
(sim.py:397-420), limiters (nebula3d.py:40-51, sim.py:366-394), max_speed (nebula3d.py:154-162), sample_n
(nebula3d.py:195-226), box_range (229-244), cast_rays (247-270), apply_front (281-307), smooth_xi (310-315),
cast_fuv (320-352), apply_fuv (355-388).

NumPy float32 arithmetic on float32 arrays is IEEE single precision element by element, so with the same
expression order this reproduces the C++ float oracle bit for bit (up to the sign of zero); float64 likewise.
Transcendentals are passed in (the oracle's deterministic float functions, or libm through `math` for double).
"""
import math

import numpy as np


def _minmod(a, b):
    return np.where(a * b <= 0, np.zeros_like(a), np.where(np.abs(a) < np.abs(b), a, b))


def _vanleer(a, b):
    with np.errstate(divide="ignore", invalid="ignore"):
        v = 2 * a * b / (a + b)
    return np.where(a * b <= 0, np.zeros_like(a), v)


def _mc(a, b):
    s = np.where(a > 0, np.ones_like(a), -np.ones_like(a))
    v = s * np.minimum(np.minimum(2 * np.abs(a), 2 * np.abs(b)), 0.5 * np.abs(a + b))
    return np.where(a * b <= 0, np.zeros_like(a), v)


LIMITERS = {0: _minmod, 1: _vanleer, 2: _mc}


def _flux(qa, qb, qc, qd, c2b, c2c, axis, floor, solver, lim):
    """qa..qd: lists of 4 primitive arrays; returns 4 flux arrays on the b|c face."""
    sl = LIMITERS[lim if solver else 0]
    L = [qb[c] + 0.5 * sl(qb[c] - qa[c], qc[c] - qb[c]) for c in range(4)]
    R = [qc[c] - 0.5 * sl(qc[c] - qb[c], qd[c] - qc[c]) for c in range(4)]
    rL = np.maximum(L[0], floor)
    rR = np.maximum(R[0], floor)
    unL, unR = L[axis + 1], R[axis + 1]
    UL = [rL, rL * L[1], rL * L[2], rL * L[3]]
    UR = [rR, rR * R[1], rR * R[2], rR * R[3]]
    FL = [UL[c] * unL for c in range(4)]
    FR = [UR[c] * unR for c in range(4)]
    FL[axis + 1] = FL[axis + 1] + rL * c2b
    FR[axis + 1] = FR[axis + 1] + rR * c2c
    if not solver:
        amax = np.maximum(np.abs(unL) + np.sqrt(c2b), np.abs(unR) + np.sqrt(c2c))
        ha = 0.5 * amax
        return [0.5 * (FL[c] + FR[c]) - ha * (UR[c] - UL[c]) for c in range(4)]
    cb, cc = np.sqrt(c2b), np.sqrt(c2c)
    SL = np.minimum(unL - cb, unR - cc)
    SR = np.maximum(unL + cb, unR + cc)
    ss, den = SL * SR, SR - SL
    out = []
    with np.errstate(divide="ignore", invalid="ignore"):
        for c in range(4):
            mid = ((SR * FL[c] - SL * FR[c]) + ss * (UR[c] - UL[c])) / den
            out.append(np.where(SL >= 0, FL[c], np.where(SR <= 0, FR[c], mid)))
    return out


def hydro_stage(U, xi, U0, stars, nstars, wind, jets, jet_cos, jet_r, pulse, dt, dx, a, b, floor, cn2, ci2, vmax,
                solver, lim):
    """One RK stage over the whole grid. U, U0: 4 arrays [nx,ny,nz]; returns the 4 output arrays.
    All scalars must already be of the array dtype."""
    dtp = U[0].dtype.type
    nx, ny, nz = U[0].shape
    pads = [np.pad(f, 2, mode="edge") for f in U]
    xpad = np.pad(xi, 1, mode="edge")
    r_p = np.maximum(pads[0], floor)
    Q = [r_p, pads[1] / r_p, pads[2] / r_p, pads[3] / r_p]
    C2 = cn2 + xpad * (ci2 - cn2)

    def q_at(off):
        return [q[2 + off[0]:2 + off[0] + nx, 2 + off[1]:2 + off[1] + ny, 2 + off[2]:2 + off[2] + nz] for q in Q]

    def c_at(off):
        return C2[1 + off[0]:1 + off[0] + nx, 1 + off[1]:1 + off[1] + ny, 1 + off[2]:1 + off[2] + nz]

    div = [np.zeros_like(U[0]) for _ in range(4)]
    for axis in range(3):
        e = [0, 0, 0]
        e[axis] = 1
        sh = lambda s: tuple(s * v for v in e)
        qm2, qm1, q0, qp1, qp2 = (q_at(sh(s)) for s in (-2, -1, 0, 1, 2))
        cm1, c0, cp1 = (c_at(sh(s)) for s in (-1, 0, 1))
        fR = _flux(qm1, q0, qp1, qp2, c0, cp1, axis, floor, solver, lim)
        fL = _flux(qm2, qm1, q0, qp1, cm1, c0, axis, floor, solver, lim)
        div = [div[c] + (fR[c] - fL[c]) for c in range(4)]
    dU = [-div[c] / dx for c in range(4)]
    idx = np.arange(nx, dtype=dtp), np.arange(ny, dtype=dtp), np.arange(nz, dtype=dtp)
    p = [((idx[0] + dtp(0.5)) * dx)[:, None, None], ((idx[1] + dtp(0.5)) * dx)[None, :, None],
         ((idx[2] + dtp(0.5)) * dx)[None, None, :]]
    if wind > 0:
        sel = xi > dtp(0.01)
        acc = [np.zeros_like(U[0]) for _ in range(3)]
        for s in range(nstars):
            st = stars[s]
            if st[3] > 0:
                rel = [p[c] - st[c] for c in range(3)]
                r2 = ((rel[0] * rel[0] + rel[1] * rel[1]) + rel[2] * rel[2]) + dtp(0.0004)
                f = st[3] / (r2 * np.sqrt(r2))
                acc = [acc[c] + rel[c] * f for c in range(3)]
        wx = wind * xi
        for c in range(3):
            dU[c + 1] = np.where(sel, dU[c + 1] + U[0] * (acc[c] * wx), dU[c + 1])
    accj = [np.zeros_like(U[0]) for _ in range(3)]
    for s in range(nstars):
        jt, st = jets[s], stars[s]
        if jt[3] > 0 and st[3] > 0:
            rel = [np.broadcast_to(p[c] - st[c], U[0].shape) for c in range(3)]
            r = np.sqrt((rel[0] * rel[0] + rel[1] * rel[1]) + rel[2] * rel[2])
            with np.errstate(divide="ignore", invalid="ignore"):
                cth = ((rel[0] * jt[0] + rel[1] * jt[1]) + rel[2] * jt[2]) / r
                f = np.where(cth > 0, dtp(1), dtp(-1)) * jt[3] * pulse / (r * r + dtp(0.0004))
            m = (r < jet_r) & (r > dtp(1e-6)) & (np.abs(cth) > jet_cos)
            accj = [np.where(m, accj[c] + jt[c] * f, accj[c]) for c in range(3)]
    for c in range(3):
        dU[c + 1] = dU[c + 1] + U[0] * accj[c]
    out = [a * U0[c] + b * (U[c] + dt * dU[c]) for c in range(4)]
    r = np.maximum(out[0], floor)
    vel = [out[c + 1] / r for c in range(3)]
    sp = np.sqrt((vel[0] * vel[0] + vel[1] * vel[1]) + vel[2] * vel[2])
    with np.errstate(divide="ignore", invalid="ignore"):
        f = vmax / sp
    cap = sp > vmax
    vel = [np.where(cap, v * f, v) for v in vel]
    return [r, r * vel[0], r * vel[1], r * vel[2]]


def max_speed(rho, mx, my, mz, xi, cn2, ci2):
    s = ((np.abs(mx / rho) + np.abs(my / rho)) + np.abs(mz / rho)) + np.sqrt(cn2 + xi * (ci2 - cn2))
    return s.max()


def sample_n(f, p, dx):
    """trilinear samples of field f at points p (N, 3)."""
    dtp = f.dtype.type
    nx, ny, nz = f.shape
    g = p / dx - dtp(0.5)
    out = np.zeros(len(p), f.dtype)
    ok = ~((g < -0.5).any(axis=1) | (g[:, 0] > dtp(nx) - dtp(0.5)) | (g[:, 1] > dtp(ny) - dtp(0.5))
           | (g[:, 2] > dtp(nz) - dtp(0.5)))
    g = g[ok]
    i0 = np.floor(g).astype(np.int64)
    fr = g - i0.astype(f.dtype)
    n = np.array([nx, ny, nz])
    lo = np.clip(i0, 0, n - 1)
    hi = np.clip(i0 + 1, 0, n - 1)
    c = {}
    for a in (0, 1):
        for b in (0, 1):
            for cc in (0, 1):
                ii = (hi if a else lo)[:, 0], (hi if b else lo)[:, 1], (hi if cc else lo)[:, 2]
                c[a, b, cc] = f[ii]
    c00 = c[0, 0, 0] + (c[1, 0, 0] - c[0, 0, 0]) * fr[:, 0]
    c10 = c[0, 1, 0] + (c[1, 1, 0] - c[0, 1, 0]) * fr[:, 0]
    c01 = c[0, 0, 1] + (c[1, 0, 1] - c[0, 0, 1]) * fr[:, 0]
    c11 = c[0, 1, 1] + (c[1, 1, 1] - c[0, 1, 1]) * fr[:, 0]
    c0 = c00 + (c10 - c00) * fr[:, 1]
    c1 = c01 + (c11 - c01) * fr[:, 1]
    out[ok] = c0 + (c1 - c0) * fr[:, 2]
    return out


def ray_dirs(nu, nphi, dtp, sincos):
    pi = dtp(math.pi)
    iu = np.arange(nu, dtype=dtp)[:, None]
    ip = np.arange(nphi, dtype=dtp)[None, :]
    u = dtp(-1) + (iu + dtp(0.5)) * dtp(2) / dtp(nu)
    s = np.sqrt(np.maximum(dtp(1) - u * u, dtp(0)))
    phi = (ip + dtp(0.5)) * dtp(2) * pi / dtp(nphi)
    sn = np.array([sincos(v)[0] for v in phi.ravel()], dtp).reshape(phi.shape)
    cs = np.array([sincos(v)[1] for v in phi.ravel()], dtp).reshape(phi.shape)
    d = np.stack(np.broadcast_arrays(s * cs, s * sn, u * np.ones_like(phi)), axis=-1)
    return d.reshape(-1, 3)


def box_range(o, d, bmax):
    """vectorised slab test for directions d (N,3)."""
    dtp = d.dtype.type
    t0 = np.zeros(len(d), d.dtype)
    t1 = np.full(len(d), dtp(1e9))
    for a in range(3):
        da = d[:, a]
        small = np.abs(da) < dtp(1e-9)
        with np.errstate(divide="ignore", invalid="ignore"):
            ta = (dtp(0) - dtp(o[a])) / da
            tb = (dtp(bmax[a]) - dtp(o[a])) / da
        miss = small & ((dtp(o[a]) < 0) | (dtp(o[a]) > dtp(bmax[a])))
        t0 = np.where(small, t0, np.maximum(t0, np.minimum(ta, tb)))
        t1 = np.where(small, np.where(miss, dtp(-1), t1), np.minimum(t1, np.maximum(ta, tb)))
    return t0, t1


def cast_rays(rho, star, Q, dr, nsteps, dx, rcap, bmax, nu, nphi, sincos):
    dtp = rho.dtype.type
    d = ray_dirs(nu, nphi, dtp, sincos)
    st = np.array(star, dtp)
    t0, t1 = box_range(st, d, bmax)
    n0 = (t0 / dtp(dr)).astype(np.int64)
    n1 = np.minimum(nsteps, (t1 / dtp(dr)).astype(np.int64) + 2)
    col = np.zeros(len(d), rho.dtype)
    rf = np.full(len(d), dtp(rcap))
    live = np.ones(len(d), bool)
    for n in range(int(n0.min()), int(n1.max())):
        act = live & (n >= n0) & (n < n1)
        if not act.any():
            continue
        r = (dtp(n) + dtp(0.5)) * dtp(dr)
        nv = sample_n(rho, st[None, :] + d[act] * r, dtp(dx))
        inc = nv * nv * r * r * dtp(dr)
        hit = col[act] + inc > dtp(Q)
        idx = np.flatnonzero(act)
        rf[idx[hit]] = r - dtp(0.5) * dtp(dr) + dtp(dr) * (dtp(Q) - col[idx[hit]]) / np.maximum(inc[hit], dtp(1e-20))
        live[idx[hit]] = False
        col[idx[~hit]] = col[idx[~hit]] + inc[~hit]
    return rf.reshape(nu, nphi)


def cell_sph(shape, star, dx, atan2):
    dtp = type(dx)
    nx, ny, nz = shape
    idx = [np.arange(m, dtype=dtp) for m in shape]
    rel = [np.broadcast_to(((idx[0] + dtp(0.5)) * dx)[:, None, None] - dtp(star[0]), shape),
           np.broadcast_to(((idx[1] + dtp(0.5)) * dx)[None, :, None] - dtp(star[1]), shape),
           np.broadcast_to(((idx[2] + dtp(0.5)) * dx)[None, None, :] - dtp(star[2]), shape)]
    r = np.maximum(np.sqrt((rel[0] * rel[0] + rel[1] * rel[1]) + rel[2] * rel[2]), dtp(1e-6))
    u = rel[2] / r
    phi = np.array([atan2(y, x) for y, x in zip(rel[1].ravel(), rel[0].ravel())], dtp).reshape(shape)
    phi = np.where(phi < 0, phi + dtp(2) * dtp(math.pi), phi)
    return r, u, phi


def apply_front(rfront, star, dx, width, xi, df, atan2):
    dtp = xi.dtype.type
    nu, nphi = rfront.shape
    r, u, phi = cell_sph(xi.shape, star, dtp(dx), atan2)
    fu = (u + dtp(1)) * dtp(0.5) * dtp(nu) - dtp(0.5)
    fp = phi / (dtp(2) * dtp(math.pi)) * dtp(nphi) - dtp(0.5)
    iu = np.floor(fu).astype(np.int64)
    ip = np.floor(fp).astype(np.int64)
    tu = fu - iu.astype(dtp)
    tp = fp - ip.astype(dtp)
    at = lambda a, b: rfront[np.clip(a, 0, nu - 1), np.mod(b, nphi)]
    av = at(iu, ip) * (dtp(1) - tp) + at(iu, ip + 1) * tp
    bv = at(iu + 1, ip) * (dtp(1) - tp) + at(iu + 1, ip + 1) * tp
    rf = av * (dtp(1) - tu) + bv * tu
    val = np.minimum(np.maximum((rf - r) / dtp(width) + dtp(0.5), dtp(0)), dtp(1))
    xi[...] = np.where(val > xi, val, xi)
    df[...] = np.minimum(df, r - rf)


def smooth_xi(src):
    p = np.pad(src, 1, mode="edge")
    nx, ny, nz = src.shape
    s = lambda a, b, c: p[1 + a:1 + a + nx, 1 + b:1 + b + ny, 1 + c:1 + c + nz]
    m = (s(-1, 0, 0) + s(1, 0, 0) + s(0, -1, 0) + s(0, 1, 0) + s(0, 0, -1) + s(0, 0, 1)) / src.dtype.type(6)
    return src.dtype.type(0.5) * src + src.dtype.type(0.5) * m
