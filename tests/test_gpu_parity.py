"""GPU parity: every C-ABI entry point and the whole Sim step against the CPU oracle, bit for bit.

Bitwise equality (numpy array_equal, so -0 == +0) is the bar: the kernels use no FMA contraction, IEEE
division and square root, and the same deterministic transcendentals as the oracle. Float comparisons of
the sign of zero are excluded because no divisor in the step can be zero (floor > 0, c^2 >= cn2 > 0, SR > SL
in the HLL branch that divides).
"""
import ctypes
import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F32 = np.float32
f = ctypes.c_float


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def env():
    import torch
    from nebulaviewer_amd import _lib
    _lib.load()
    dev = torch.device("cuda:0")
    return torch, dev, _lib


def T(env, a):
    torch, dev, _ = env
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def N(t):
    return t.detach().cpu().numpy()


def stream(env):
    torch, dev, _ = env
    return torch.cuda.current_stream(dev).cuda_stream


def random_state(shape, seed):
    rng = np.random.default_rng(seed)
    rho = rng.lognormal(0.0, 1.0, shape).clip(0.001, None).astype(F32)
    m = [(rho * rng.normal(0, 1.0, shape)).astype(F32) for _ in range(3)]
    fast = rng.uniform(size=shape) < 0.05
    m[0][fast] = rho[fast] * F32(9.0)
    rho[rng.uniform(size=shape) < 0.03] = F32(1e-4)
    xi = rng.uniform(0, 1, shape).astype(F32)
    xi[rng.uniform(size=shape) < 0.3] = 0
    return [rho, *m], xi


def tables(shape, dx, nst=3):
    L = np.array(shape) * dx
    stars = np.zeros((max(nst, 1), 4), F32)
    jets = np.zeros((max(nst, 1), 4), F32)
    base = [[0.3, 0.6, 0.5, 1.0], [0.7, 0.2, 0.4, 0.5], [0.5, 0.5, 0.5, 0.0]]
    for s in range(nst):
        stars[s] = [base[s][0] * L[0], base[s][1] * L[1], base[s][2] * L[2], base[s][3]]
    ax = np.array([0.3, -0.4, 0.866])
    ax /= np.linalg.norm(ax)
    jb = [[*ax, 2.0], [0, 0, 1, 0.0], [1, 0, 0, 3.0]]
    for s in range(nst):
        jets[s] = jb[s]
    return stars, jets


SHAPES = [(12, 9, 14), (70, 33, 130), (5, 3, 2), (1, 1, 64), (40, 64, 64)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("solver,lim", [(1, 0), (1, 1), (1, 2), (0, 0)])
@pytest.mark.parametrize("mode", ["stage1", "stage2", "inplace"])
def test_hydro_stage_bitwise(env, okern32, shape, solver, lim, mode):
    say(f"[hydro] shape={shape} solver={solver} lim={lim} mode={mode}")
    torch, dev, L = env
    dx = F32(1.0 / max(shape))
    U, xi = random_state(shape, seed=hash((shape, solver, lim)) % 1000)
    stars, jets = tables(shape, float(dx))
    U0n = U if mode == "stage1" else random_state(shape, seed=7)[0]
    a, b = (0.0, 1.0) if mode == "stage1" else (0.5, 0.5)
    sc = dict(nstars=3, wind=0.7, jet_cos=math.cos(math.radians(30)), jet_r=0.6, pulse=1.3, dt=0.002, dx=float(dx),
              a=a, b=b, floor=2e-3, cn2=(1 / 12) ** 2, ci2=1.0, vmax=6.0, solver=solver, lim=lim)
    ref_out = [np.zeros(shape, F32) for _ in range(4)]
    okern32.hydro_stage(U, xi, U0n, ref_out, shape, stars, jets=jets, **sc)
    ref_smax = okern32.max_speed(*ref_out, xi, sc["cn2"], sc["ci2"])

    gU = [T(env, u) for u in U]
    gxi = T(env, xi)
    gU0 = gU if mode == "stage1" else [T(env, u) for u in U0n]
    gout = gU0 if mode == "inplace" else [torch.zeros(shape, dtype=torch.float32, device=dev) for _ in range(4)]
    smax = torch.zeros(1, dtype=torch.float32, device=dev)
    gst, gjt = T(env, stars), T(env, jets)
    L.call("nv_hydro_stage", *[t.data_ptr() for t in gU], gxi.data_ptr(), *[t.data_ptr() for t in gU0],
           *[t.data_ptr() for t in gout], *shape, gst.data_ptr(), 3, f(sc["wind"]), gjt.data_ptr(), f(sc["jet_cos"]),
           f(sc["jet_r"]), f(sc["pulse"]), f(sc["dt"]), f(sc["dx"]), f(a), f(b), f(sc["floor"]), f(sc["cn2"]),
           f(sc["ci2"]), f(sc["vmax"]), solver, lim, smax.data_ptr(), stream(env))
    torch.cuda.synchronize()
    for c in range(4):
        np.testing.assert_array_equal(N(gout[c]), ref_out[c], err_msg=f"field {c}")
    assert float(smax.item()) == ref_smax


# (40, 48, 192) with one-plane x chunks has more tiles than the device holds workgroups at once (720 of 64 x 8, 360
# of 64 x 16): those launches deal the tiles to the XCDs in bands
@pytest.mark.parametrize("shape", [(12, 9, 14), (33, 40, 96), (9, 35, 160), (17, 24, 64), (40, 48, 192)])
# shapes 1-3, and shapes with a march depth (NV_TILE_DEPTH: one-plane and three-plane x chunks, whole columns)
@pytest.mark.parametrize("tile", [1, 2, 3, 1 | (1 << 8), 2 | (3 << 8), 3 | (1 << 8), 3 | (1 << 23), 1 | (1 << 15)])
@pytest.mark.parametrize("precision", [0, 2])
def test_hydro_stage_tile_shapes(env, okern32, shape, tile, precision):
    """nv_hydro_stage_ex with each workgroup tile shape and march depth: bitwise the oracle (FP32 exact), bitwise
    the automatic shape's result (binary16-velocity form) - shape and depth change only the work decomposition"""
    say(f"[hydro tile] shape={shape} tile={tile} precision={precision}")
    torch, dev, L = env
    dx = F32(1.0 / max(shape))
    U, xi = random_state(shape, seed=5)
    stars, jets = tables(shape, float(dx))
    U0n = random_state(shape, seed=6)[0]
    gst, gjt = T(env, stars), T(env, jets)
    gxi = T(env, xi)
    if precision == 0:
        gU, gU0 = [T(env, u) for u in U], [T(env, u) for u in U0n]
    else:
        def enc(V):
            # the binary16-velocity state holds |v| <= vmax (the stage caps it): scale faster cells down to 5.9
            r = np.maximum(V[0], F32(2e-3))
            v = np.stack([V[c] / r for c in (1, 2, 3)])
            sp = np.sqrt((v.astype(np.float64) ** 2).sum(0))
            v = (v * np.minimum(1.0, 5.9 / np.maximum(sp, 1e-30))).astype(np.float32)
            assert np.isfinite(v.astype(np.float16)).all()
            return [T(env, V[0])] + [T(env, v[c].astype(np.float16)) for c in range(3)]
        gU, gU0 = enc(U), enc(U0n)
    outs = {}
    for t in (tile, 0):
        out = [torch.zeros_like(g) for g in gU]
        smax = torch.zeros(1, dtype=torch.float32, device=dev)
        d = L.NvStageDesc()
        for c in range(4):
            d.u[c], d.u0[c], d.out[c] = gU[c].data_ptr(), gU0[c].data_ptr(), out[c].data_ptr()
        d.xi, d.stars, d.jets, d.smax_out, d.dt_dev = gxi.data_ptr(), gst.data_ptr(), gjt.data_ptr(), smax.data_ptr(), None
        d.nx, d.ny, d.nz = shape
        d.nstars, d.wind, d.jet_cos, d.jet_r, d.pulse, d.dt, d.dx = 3, 0.7, math.cos(math.radians(30)), 0.6, 1.3, 0.002, float(dx)
        d.a, d.b, d.floor_, d.cn2, d.ci2, d.vmax = 0.5, 0.5, 2e-3, (1 / 12) ** 2, 1.0, 6.0
        L.call("nv_hydro_stage_ex", ctypes.byref(d), precision, 1, 0, t, stream(env))
        torch.cuda.synchronize()
        outs[t] = [N(o) for o in out] + [float(smax.item())]
    for c in range(5):
        np.testing.assert_array_equal(outs[tile][c], outs[0][c], err_msg=f"field {c}")
    if precision == 0:
        ref_out = [np.zeros(shape, F32) for _ in range(4)]
        okern32.hydro_stage(U, xi, U0n, ref_out, shape, stars, 3, 0.7, jets, math.cos(math.radians(30)), 0.6, 1.3,
                            0.002, float(dx), 0.5, 0.5, 2e-3, (1 / 12) ** 2, 1.0, 6.0, 1, 0)
        for c in range(4):
            np.testing.assert_array_equal(outs[tile][c], ref_out[c], err_msg=f"field {c}")


def tie_state(shape, seed):
    """momenta whose quotient by the density is an exact tie of the subnormal grid: rho an even integer B,
    m = +-B M 2^-150 with M odd, so m / rho = M 2^-150 lies halfway between two subnormal floats (IEEE rounds to
    the even one); a fifth of the cells hold ordinary momenta"""
    rng = np.random.default_rng(seed)
    B = (rng.integers(2, 101, shape) * 2).astype(np.int64)
    U = [B.astype(F32)]
    for _ in range(3):
        M = rng.integers(0, 1024, shape) * 2 + 1
        m = np.ldexp((B * M).astype(np.float64), -150).astype(F32)
        assert np.array_equal(m.astype(np.float64), np.ldexp((B * M).astype(np.float64), -150))
        m = np.where(rng.uniform(size=shape) < 0.5, -m, m).astype(F32)
        big = rng.uniform(size=shape) < 0.2
        m[big] = (U[0][big] * rng.normal(0, 0.5, int(big.sum()))).astype(F32)
        U.append(m)
    xi = rng.uniform(0, 1, shape).astype(F32)
    return U, xi


@pytest.mark.parametrize("tile", [1, 2, 3])
@pytest.mark.parametrize("solver", [1, 0])
def test_hydro_stage_subnormal_ties(env, okern32, tile, solver):
    """quotients on exact ties of the subnormal grid (decoded velocities, and the fluxes and updated velocities
    they feed) come out as the oracle's IEEE divisions, bit for bit"""
    say(f"[hydro ties] tile={tile} solver={solver}")
    torch, dev, L = env
    shape = (10, 20, 64)
    dx = 1.0 / 64
    U, xi = tie_state(shape, 21)
    U0n, _ = tie_state(shape, 22)
    q = (U[1] / U[0]).astype(F32)          # numpy's IEEE quotient: the ties went to even neighbours
    assert (np.abs(q[q != 0]) < np.finfo(F32).tiny).mean() > 0.5
    gU, gU0, gxi = [T(env, u) for u in U], [T(env, u) for u in U0n], T(env, xi)
    out = [torch.zeros(shape, dtype=torch.float32, device=dev) for _ in range(4)]
    smax = torch.zeros(1, dtype=torch.float32, device=dev)
    d = L.NvStageDesc()
    for c in range(4):
        d.u[c], d.u0[c], d.out[c] = gU[c].data_ptr(), gU0[c].data_ptr(), out[c].data_ptr()
    d.xi, d.stars, d.jets, d.smax_out, d.dt_dev = gxi.data_ptr(), None, None, smax.data_ptr(), None
    d.nx, d.ny, d.nz = shape
    d.nstars, d.wind, d.jet_cos, d.jet_r, d.pulse, d.dt, d.dx = 0, 0.0, 0.9, 0.04, 1.0, 1e-4, dx
    d.a, d.b, d.floor_, d.cn2, d.ci2, d.vmax = 0.5, 0.5, 2e-3, (1 / 12) ** 2, 1.0, 6.0
    L.call("nv_hydro_stage_ex", ctypes.byref(d), 0, solver, 0, tile, stream(env))
    torch.cuda.synchronize()
    ref = [np.zeros(shape, F32) for _ in range(4)]
    okern32.hydro_stage(U, xi, U0n, ref, shape, np.zeros((1, 4), F32), 0, 0.0, np.zeros((1, 4), F32), 0.9, 0.04, 1.0,
                        1e-4, dx, 0.5, 0.5, 2e-3, (1 / 12) ** 2, 1.0, 6.0, solver, 0)
    for c in range(4):
        np.testing.assert_array_equal(N(out[c]), ref[c], err_msg=f"field {c}")
    assert float(smax.item()) == okern32.max_speed(*ref, xi, (1 / 12) ** 2, 1.0)


def test_hydro_stage_no_stars(env, okern32):
    say("[hydro] no stars, null tables")
    torch, dev, L = env
    shape = (33, 17, 70)
    U, xi = random_state(shape, 3)
    ref = [np.zeros(shape, F32) for _ in range(4)]
    okern32.hydro_stage(U, xi, U, ref, shape, np.zeros((1, 4), F32), 0, 0.0, np.zeros((1, 4), F32), 0.9, 0.04, 1.0,
                        0.001, 1 / 70, 0.0, 1.0, 2e-3, 1 / 144, 1.0, 6.0, 1, 0)
    gU = [T(env, u) for u in U]
    gxi = T(env, xi)
    gout = [torch.zeros(shape, dtype=torch.float32, device=dev) for _ in range(4)]
    L.call("nv_hydro_stage", *[t.data_ptr() for t in gU], gxi.data_ptr(), *[t.data_ptr() for t in gU],
           *[t.data_ptr() for t in gout], *shape, None, 0, f(0), None, f(0.9), f(0.04), f(1), f(0.001), f(1 / 70),
           f(0), f(1), f(2e-3), f(1 / 144), f(1), f(6), 1, 0, None, stream(env))
    torch.cuda.synchronize()
    for c in range(4):
        np.testing.assert_array_equal(N(gout[c]), ref[c])


@pytest.mark.parametrize("shape", [(12, 9, 14), (64, 64, 64), (3, 200, 7)])
def test_max_speed_bitwise(env, okern32, shape):
    say(f"[max_speed] {shape}")
    torch, dev, L = env
    U, xi = random_state(shape, 4)
    U[0] = np.maximum(U[0], F32(2e-3))
    out = torch.zeros(1, dtype=torch.float32, device=dev)
    g = [T(env, u) for u in U] + [T(env, xi)]
    L.call("nv_max_speed", *[t.data_ptr() for t in g], *shape, f(1 / 144), f(1.0), out.data_ptr(), stream(env))
    assert float(out.item()) == okern32.max_speed(*U, xi, 1 / 144, 1.0)


def ray_case(shape, seed=11):
    rng = np.random.default_rng(seed)
    rho = rng.lognormal(0, 1.0, shape).astype(F32)
    dx = 1.0 / shape[2]
    Lb = tuple(n * dx for n in shape)
    diag = math.sqrt(sum(v * v for v in Lb)) + 0.6
    return rho, dx, Lb, diag


@pytest.mark.parametrize("star,Q", [((0.4, 0.35, 0.5), 0.01), ((0.95, 0.05, 0.1), 0.05), ((-0.2, 0.4, 0.5), 0.003),
                                    ((0.5, 0.5, 0.5), 10.0)])
def test_cast_rays_bitwise(env, okern32, star, Q):
    say(f"[cast_rays] star={star} Q={Q}")
    torch, dev, L = env
    shape = (40, 36, 44)
    rho, dx, Lb, diag = ray_case(shape)
    dr = 0.5 * dx
    nsteps = int(diag / dr) + 1
    nu, nphi = 32, 64
    ref = np.zeros((nu, nphi), F32)
    okern32.cast_rays(rho, star, Q, dr, nsteps, dx, diag, Lb, ref)
    g = torch.zeros((nu, nphi), dtype=torch.float32, device=dev)
    grho = T(env, rho)
    L.call("nv_cast_rays", grho.data_ptr(), *shape, *(f(v) for v in star), f(Q), f(dr), nsteps, f(dx), f(diag),
           *(f(v) for v in Lb), g.data_ptr(), nu, nphi, stream(env))
    np.testing.assert_array_equal(N(g), ref)


@pytest.mark.parametrize("shape,dx", [((32, 16, 64), 1 / 64), ((20, 24, 1), 1 / 24), ((20, 24, 1), 1 / 32),
                                      ((9, 7, 2), 0.1)])
def test_cast_rays_sampler_forms_bitwise(env, okern32, shape, dx):
    """power-of-two dx (multiply by 1/dx), general dx (shared reciprocal), nz = 1 (no paired z loads), nz = 2"""
    say(f"[cast_rays forms] shape={shape} dx={dx}")
    torch, dev, L = env
    rho = np.random.default_rng(31).lognormal(0, 1.0, shape).astype(F32)
    Lb = tuple(n * dx for n in shape)
    diag = math.sqrt(sum(v * v for v in Lb)) + 0.6
    dr = 0.5 * dx
    nsteps = int(diag / dr) + 1
    star = (0.37 * Lb[0], 0.52 * Lb[1], 0.45 * Lb[2])
    nu, nphi = 24, 40
    grho = T(env, rho)
    for Q in (1e-4, 3e-3, 1e3):
        ref = np.zeros((nu, nphi), F32)
        okern32.cast_rays(rho, star, Q, dr, nsteps, dx, diag, Lb, ref)
        g = torch.zeros((nu, nphi), dtype=torch.float32, device=dev)
        L.call("nv_cast_rays", grho.data_ptr(), *shape, *(f(v) for v in star), f(Q), f(dr), nsteps, f(dx),
               f(diag), *(f(v) for v in Lb), g.data_ptr(), nu, nphi, stream(env))
        np.testing.assert_array_equal(N(g), ref)


def bricks_numpy(rho):
    """the ghost-layered bricks: brick (p, q, r) = (rho(i, ja, ka), rho(i, ja, kb), rho(i, jb, ka), rho(i, jb, kb)),
    i = clamp(p-1), ja = clamp(q-1), jb = clamp(q), ka = clamp(r-1), kb = clamp(r); shape (nx+2, ny+1, nz+1, 4)"""
    nx, ny, nz = rho.shape
    i = np.clip(np.arange(nx + 2) - 1, 0, nx - 1)
    ja, jb = np.clip(np.arange(ny + 1) - 1, 0, ny - 1), np.clip(np.arange(ny + 1), 0, ny - 1)
    ka, kb = np.clip(np.arange(nz + 1) - 1, 0, nz - 1), np.clip(np.arange(nz + 1), 0, nz - 1)
    r = rho[i]
    return np.stack([r[:, ja][:, :, ka], r[:, ja][:, :, kb], r[:, jb][:, :, ka], r[:, jb][:, :, kb]], axis=-1)


@pytest.mark.parametrize("shape,dx", [((40, 36, 44), 1 / 44), ((32, 16, 64), 1 / 64), ((20, 24, 1), 1 / 24),
                                      ((9, 7, 2), 0.1), ((7, 1, 30), 1 / 30), ((1, 1, 1), 0.5)])
def test_cast_rays_bricks_bitwise(env, okern32, shape, dx):
    """nv_ray_bricks against its numpy construction, and nv_cast_rays_ex sampling the bricks against the oracle (float
    and FP64-column forms), with stars inside, at the faces and outside the box; ny = 1, nz = 1 and a single cell
    exercise every clamped corner"""
    say(f"[cast_rays bricks] shape={shape} dx={dx}")
    torch, dev, L = env
    rho = np.random.default_rng(41).lognormal(0, 1.0, shape).astype(F32)
    grho = T(env, rho)
    pshape = (shape[0] + 2, shape[1] + 1, shape[2] + 1, 4)
    assert L.load().nv_ray_bricks_size(*shape) == 4 * math.prod(pshape)
    gb = torch.full(pshape, -1.0, dtype=torch.float32, device=dev)
    L.call("nv_ray_bricks", grho.data_ptr(), *shape, gb.data_ptr(), stream(env))
    np.testing.assert_array_equal(N(gb), bricks_numpy(rho))
    Lb = tuple(n * dx for n in shape)
    diag = math.sqrt(sum(v * v for v in Lb)) + 0.6
    dr = 0.5 * dx
    nsteps = int(diag / dr) + 1
    nu, nphi = 24, 40
    for star in ((0.37 * Lb[0], 0.52 * Lb[1], 0.45 * Lb[2]), (0.0, Lb[1], 0.5 * Lb[2]), (-0.3, 0.4, 1.2 * Lb[2])):
        for Q in (1e-4, 3e-3, 1e3):
            for wide in (0, 1):
                ref = np.zeros((nu, nphi), F32)
                okern32.cast_rays(rho, star, Q, dr, nsteps, dx, diag, Lb, ref, wide=bool(wide))
                g = torch.zeros((nu, nphi), dtype=torch.float32, device=dev)
                L.call("nv_cast_rays_ex", grho.data_ptr(), gb.data_ptr(), *shape, *(f(v) for v in star), f(Q), f(dr),
                       nsteps, f(dx), f(diag), *(f(v) for v in Lb), g.data_ptr(), nu, nphi, wide, stream(env))
                np.testing.assert_array_equal(N(g), ref, err_msg=f"star {star} Q {Q} wide {wide}")
    # bricks = null: the paired sampler, the same entry point
    ref = np.zeros((nu, nphi), F32)
    okern32.cast_rays(rho, star, 3e-3, dr, nsteps, dx, diag, Lb, ref)
    g = torch.zeros((nu, nphi), dtype=torch.float32, device=dev)
    L.call("nv_cast_rays_ex", grho.data_ptr(), None, *shape, *(f(v) for v in star), f(3e-3), f(dr), nsteps, f(dx),
           f(diag), *(f(v) for v in Lb), g.data_ptr(), nu, nphi, 0, stream(env))
    np.testing.assert_array_equal(N(g), ref)


def test_cast_rays_wide_bitwise(env, okern32, okern64):
    """The FP64 long-ray column against its oracle on the stalled-column case, and the float form beside it."""
    import long_ray_case as LR
    say("[cast_rays_wide] stalled-column case 640x24x24")
    torch, dev, L = env
    Q = LR.find_q(okern64)
    rho, star, dr, nsteps, dx, diag, Lb = LR.geometry()
    grho = T(env, rho)
    for entry, wide in (("nv_cast_rays_wide", True), ("nv_cast_rays", False)):
        ref = np.zeros(LR.RAYS, F32)
        okern32.cast_rays(rho, star, Q, dr, nsteps, dx, diag, Lb, ref, wide=wide)
        g = torch.zeros(LR.RAYS, dtype=torch.float32, device=dev)
        L.call(entry, grho.data_ptr(), *LR.SHAPE, *(f(v) for v in star), f(Q), f(dr), nsteps, f(dx), f(diag),
               *(f(v) for v in Lb), g.data_ptr(), *LR.RAYS, stream(env))
        np.testing.assert_array_equal(N(g), ref)
        if wide:
            wide_rf = N(g)
    assert abs(float(wide_rf[LR.RAY]) - 2 * LR.X1 * dx) < 12 * dx


def test_apply_front_smooth_half_fill_bitwise(env, okern32):
    say("[apply_front/smooth/to_half/fill]")
    torch, dev, L = env
    shape = (40, 36, 44)
    rng = np.random.default_rng(12)
    rf = rng.uniform(0.05, 0.7, (32, 64)).astype(F32)
    xi = rng.uniform(0, 0.5, shape).astype(F32)
    df = rng.uniform(-0.3, 1.0, shape).astype(F32)
    gxi, gdf = T(env, xi), T(env, df)
    star = (0.45, 0.35, 0.55)
    dx = 1.0 / 44
    okern32.apply_front(rf, star, dx, 1.2 * dx, xi, df)
    grf = T(env, rf)
    L.call("nv_apply_front", grf.data_ptr(), 32, 64, *(f(v) for v in star), f(dx), f(1.2 * dx),
           gxi.data_ptr(), gdf.data_ptr(), *shape, stream(env))
    np.testing.assert_array_equal(N(gxi), xi)
    np.testing.assert_array_equal(N(gdf), df)
    out = np.zeros(shape, F32)
    okern32.smooth_xi(xi, out)
    gout = torch.zeros(shape, dtype=torch.float32, device=dev)
    L.call("nv_smooth_xi", gxi.data_ptr(), gout.data_ptr(), *shape, stream(env))
    np.testing.assert_array_equal(N(gout), out)
    vals = np.concatenate([rng.standard_normal(200000) * 10.0 ** rng.uniform(-9, 5.5, 200000),
                           [65504, 65520, 2 ** -25, 3 * 2 ** -26, 1e-30, 7e4, -7e4, 0.0]]).astype(F32)
    h = torch.zeros(vals.size, dtype=torch.int16, device=dev)
    gvals = T(env, vals)
    L.call("nv_to_half", gvals.data_ptr(), h.data_ptr(), vals.size, stream(env))
    np.testing.assert_array_equal(N(h).view(np.uint16), okern32.to_half(vals).view(np.uint16))
    L.call("nv_fill_f32", gout.data_ptr(), f(1e3), gout.numel(), stream(env))
    assert bool((gout == 1e3).all())


@pytest.mark.parametrize("nst,half,shape,rays", [(1, True, (40, 36, 44), (32, 64)), (3, False, (23, 70, 130), (32, 64)),
                                                (0, True, (8, 8, 8), (32, 64)), (2, True, (9, 20, 40), (16, 1)),
                                                (1, False, (13, 17, 33), (8, 2))])
def test_ionize_apply_fused_bitwise(env, okern32, nst, half, shape, rays):
    """the fused pass against the reference's launch sequence; x chunks of 4 planes leave one-plane tails at nx = 9
    and 13; a single phi bin and two phi bins (every lookup at the wrap) exercise the front lookup's edge forms"""
    say(f"[ionize_apply] nstars={nst} half={half} shape={shape} rays={rays}")
    torch, dev, L = env
    rho, dx, Lb, diag = ray_case(shape, seed=13)
    U = [rho] + [(rho * np.random.default_rng(s).normal(0, 0.5, shape)).astype(F32) for s in range(3)]
    nu, nphi = rays
    dr = 0.5 * dx
    nsteps = int(diag / dr) + 1
    stars = [(0.4 * Lb[0], 0.35 * Lb[1], 0.5 * Lb[2]), (0.9 * Lb[0], 0.1 * Lb[1], 0.2 * Lb[2]),
             (0.2 * Lb[0], 0.8 * Lb[1], 0.6 * Lb[2])][:nst]
    Qs = [0.01, 0.05, 0.002]
    # oracle: the reference launch sequence (sim.py:930-943)
    xi = np.zeros(shape, F32)
    df = np.full(shape, 1e3, F32)
    rfs = np.zeros((max(nst, 1), nu, nphi), F32)
    for s, st in enumerate(stars):
        okern32.cast_rays(rho, st, Qs[s], dr, nsteps, dx, diag, Lb, rfs[s])
        okern32.apply_front(rfs[s], st, dx, 1.2 * dx, xi, df)
    tmp = np.zeros(shape, F32)
    okern32.smooth_xi(xi, tmp)
    xi = tmp
    dfo = okern32.to_half(df) if half else df
    smax_ref = okern32.max_speed(*U, xi, 1 / 144, 1.0)
    # GPU: casts + one fused pass
    grho = T(env, rho)
    gU = [grho] + [T(env, u) for u in U[1:]]
    grf = torch.zeros((max(nst, 1), nu, nphi), dtype=torch.float32, device=dev)
    for s, st in enumerate(stars):
        L.call("nv_cast_rays", grho.data_ptr(), *shape, *(f(v) for v in st), f(Qs[s]), f(dr), nsteps, f(dx),
               f(diag), *(f(v) for v in Lb), grf[s].data_ptr(), nu, nphi, stream(env))
    xyz = T(env, np.array(stars, F32).reshape(-1, 3) if nst else np.zeros((1, 3), F32))
    gxi = torch.full(shape, 7.0, dtype=torch.float32, device=dev)
    gdf = torch.zeros(shape, dtype=torch.float16 if half else torch.float32, device=dev)
    smax = torch.zeros(1, dtype=torch.float32, device=dev)
    L.call("nv_ionize_apply", grf.data_ptr(), nu, nphi, xyz.data_ptr(), nst, f(dx), f(1.2 * dx), gxi.data_ptr(),
           gdf.data_ptr() if half else None, None if half else gdf.data_ptr(), *shape,
           *[t.data_ptr() for t in gU], f(1 / 144), f(1.0), smax.data_ptr(), stream(env))
    torch.cuda.synchronize()
    np.testing.assert_array_equal(N(grf)[:nst], rfs[:nst])
    np.testing.assert_array_equal(N(gxi), xi)
    if half:
        np.testing.assert_array_equal(N(gdf).view(np.uint16), dfo.view(np.uint16))
    else:
        np.testing.assert_array_equal(N(gdf), dfo)
    assert float(smax.item()) == smax_ref


@pytest.mark.parametrize("shape", [(40, 36, 44), (32, 16, 64)])
def test_fuv_bitwise(env, okern32, shape):
    say(f"[fuv] cast_fuv / apply_fuv / fuv_apply {shape}")
    torch, dev, L = env
    rho, dx, Lb, diag = ray_case(shape, seed=14)
    xi = np.random.default_rng(15).uniform(0, 1, shape).astype(F32)
    dr = 0.5 * dx
    nsteps = int(diag / dr) + 1
    fg = (24, 16, 32)
    binr = diag / fg[0]
    stars = [(0.4 * Lb[0], 0.35 * Lb[1], 0.5 * Lb[2]), (0.9 * Lb[0], 0.1 * Lb[1], 0.2 * Lb[2])]
    lums = [1.0, 0.37]
    grho, gxi = T(env, rho), T(env, xi)
    gatt = torch.zeros((2,) + fg, dtype=torch.float32, device=dev)
    acc = np.zeros(shape, F32)
    gacc = torch.zeros(shape, dtype=torch.float32, device=dev)
    for s, st in enumerate(stars):
        att = np.zeros(fg, F32)
        okern32.cast_fuv(rho, xi, st, 2.5, dr, nsteps, dx, binr, Lb, att)
        L.call("nv_cast_fuv", grho.data_ptr(), gxi.data_ptr(), *shape, *(f(v) for v in st), f(2.5), f(dr), nsteps,
               f(dx), f(binr), *(f(v) for v in Lb), gatt[s].data_ptr(), *fg, stream(env))
        np.testing.assert_array_equal(N(gatt[s]), att)
        okern32.apply_fuv(att, st, lums[s], dx, binr, acc)
        L.call("nv_apply_fuv", gatt[s].data_ptr(), *fg, *(f(v) for v in st), f(lums[s]), f(dx), f(binr),
               gacc.data_ptr(), *shape, stream(env))
    np.testing.assert_array_equal(N(gacc), acc)
    g0 = torch.zeros(shape, dtype=torch.float16, device=dev)
    # keep the uploaded tables referenced until the kernel has run: a temporary tensor's block goes back to the
    # caching allocator at once and the next upload may reuse it before the launch reads it
    gxyz, glum = T(env, np.array(stars, F32)), T(env, np.array(lums, F32))
    L.call("nv_fuv_apply", gatt.data_ptr(), *fg, gxyz.data_ptr(), glum.data_ptr(), 2, f(dx), f(binr), g0.data_ptr(),
           None, *shape, stream(env))
    np.testing.assert_array_equal(N(g0).view(np.uint16), okern32.to_half(acc).view(np.uint16))


# ------------------------------------------------------------------ the whole Sim step against OracleSim
def run_pair(env, scenario, n, steps, rays, fuv, fused=True, params=None, ray_bricks="auto"):
    from nebulaviewer_amd import scenarios as S
    from nebulaviewer_amd.sim import Sim
    from oracle.oracle import OracleSim
    say(f"[sim] {scenario} n={n} steps={steps} rays={rays} fused={fused} ray_bricks={ray_bricks}")
    setup = S.make_setup(scenario, n, params=params)
    sim = Sim(fast=False, setup=setup, rays=rays, fuv_grid=fuv, verbose=False, fused=fused, ray_bricks=ray_bricks)
    ref = OracleSim(S.make_setup(scenario, n, params=params), S.constants(setup), rays, fuv, np.float32, threads=16)
    sim.step(steps)
    ref.step(steps)
    return sim, ref


def assert_same_state(sim, ref):
    assert sim.t == ref.t and sim.dt == ref.dt and sim.it == ref.it
    for name in ("rho", "mx", "my", "mz", "xi"):
        np.testing.assert_array_equal(N(getattr(sim, name)), getattr(ref, name), err_msg=name)
    np.testing.assert_array_equal(N(sim.dfront).view(np.uint16), ref.dfront.view(np.uint16), err_msg="dfront")


@pytest.mark.parametrize("scenario,n,steps,rays,fuv,bricks", [
    ("torus", 40, 12, (64, 128), (24, 32, 64), True),
    ("sheet", 32, 9, (32, 64), (16, 16, 32), False),
    ("wall", 48, 8, (64, 128), (24, 32, 64), True),
    ("wall", 48, 8, (64, 128), (24, 32, 64), False),
])
def test_sim_matches_oracle(env, scenario, n, steps, rays, fuv, bricks):
    sim, ref = run_pair(env, scenario, n, steps, rays, fuv, ray_bricks=bricks)
    assert sim.ray_bricks_used == bricks
    assert_same_state(sim, ref)
    sim.update_render_fields(force=True)
    ref.fuv_field()
    np.testing.assert_array_equal(N(sim.g0).view(np.uint16), ref.g0.view(np.uint16), err_msg="g0")


@pytest.mark.parametrize("scenario,n,steps,params,device_steps", [
    ("wall", 48, 8, None, "all"),
    ("sheet", 32, 7, dict(q_ramp=0.005), "some"),   # host clock while the luminosity ramps, then the device clock
    ("torus_jet", 32, 5, None, "none"),             # a pulsed jet: the pulse follows the clock, host clock throughout
])
def test_sim_device_clock_matches_host_clock(env, scenario, n, steps, params, device_steps):
    """Sim.step(k) with the clock on the device (nv_cfl_dt, no readback per step) against the host clock
    (device_clock=False: Sim._compute_dt every step): the same fields, t, dt and it bit for bit, then again with a
    dt cap, in calls of k steps against k calls of one step"""
    from nebulaviewer_amd import scenarios as S
    from nebulaviewer_amd.sim import Sim
    say(f"[sim clock] {scenario} n={n} steps={steps} params={params}")
    sims = [Sim(fast=False, setup=S.make_setup(scenario, n, params=params), rays=(32, 64), fuv_grid=(16, 16, 32),
                verbose=False, device_clock=dc) for dc in (True, False)]
    dev, host = sims
    dev.step(steps)
    host.step(steps)
    used = dev.prof.get("device_clock_steps", 0)
    assert {"all": used == steps, "some": 0 < used < steps, "none": used == 0}[device_steps], used
    assert host.prof.get("device_clock_steps", 0) == 0

    def same():
        assert dev.t == host.t and dev.dt == host.dt and dev.it == host.it, (dev.t, host.t, dev.dt, host.dt)
        for name in ("rho", "mx", "my", "mz", "xi", "dfront"):
            np.testing.assert_array_equal(N(getattr(dev, name)), N(getattr(host, name)), err_msg=name)
    same()
    for s in sims:
        s.dt_cap = 0.9 * host.dt
    dev.step(3)
    for _ in range(3):
        host.step(1)
    same()


def test_sim_unfused_sequence_matches_oracle(env):
    sim, ref = run_pair(env, "sheet", 24, 5, (32, 64), (16, 16, 32), fused=False)
    assert_same_state(sim, ref)


def test_sim_jet_and_rusanov(env):
    sim, ref = run_pair(env, "torus_jet", 32, 6, (32, 64), (16, 16, 32), params=dict(solver="rusanov"))
    assert_same_state(sim, ref)


def test_sim_sheet_256_matches_oracle(env):
    """The config-1 grid (256^3, 4 stars, wind, Mach-5 velocity), 2 steps with one ionization, bitwise."""
    sim, ref = run_pair(env, "sheet", 256, 2, (512, 1024), (160, 160, 320), ray_bricks=True)
    assert_same_state(sim, ref)


def test_sim_long_rays_matches_oracle(env, okern64):
    """Rays whose chord exceeds 512 cells (a 640 x 24 x 24 grid, the stalled-column density of
    tests/long_ray_case.py, the star at the low-x face): with long_rays="fp64" (the default) Sim keeps their column in
    FP64 and equals the oracle's FP64-column run bit for bit, with "fp32" the float-column run; the two differ."""
    sys.path.insert(0, __import__("os").path.dirname(__file__))
    import long_ray_case as LR
    from nebulaviewer_amd import scenarios as S
    from nebulaviewer_amd.sim import Sim
    from oracle.oracle import OracleSim
    rho, star, dr, nsteps, dx, diag, Lb = LR.geometry()
    Q = LR.find_q(okern64)
    P = S.make_params(dict(q_main=Q))

    def setup():
        return S.Setup("long", LR.SHAPE, LR.SHAPE[0], dict(P), rho.copy(), stars=[(*star, 1.0, 0.0)],
                       kinds=["main"], jets=[(0.0, 0.0, 1.0)], jkinds=["none"])

    xis = {}
    for mode in ("fp64", "fp32"):
        say(f"[sim] long rays {LR.SHAPE}, long_rays={mode}, 3 steps")
        sim = Sim(fast=False, setup=setup(), rays=LR.RAYS, fuv_grid=(8, 8, 16), verbose=False, long_rays=mode)
        ref = OracleSim(setup(), S.constants(setup()), LR.RAYS, (8, 8, 16), np.float32, threads=16, long_rays=mode)
        sim.step(3)
        ref.step(3)
        assert_same_state(sim, ref)
        xis[mode] = N(sim.xi)
    assert not np.array_equal(xis["fp64"], xis["fp32"])


# ------------------------------------------------------------------ size-independent properties at full size
def test_uniform_state_is_steady_512(env):
    say("[property] uniform flow at 512^3 stays bitwise unchanged")
    torch, dev, L = env
    shape = (512, 512, 512)
    kw = dict(dtype=torch.float32, device=dev)
    U = [torch.full(shape, 3.0, **kw), torch.full(shape, 0.6, **kw), torch.full(shape, -0.3, **kw),
         torch.full(shape, 0.15, **kw)]
    xi = torch.full(shape, 0.25, **kw)
    out = [torch.empty(shape, **kw) for _ in range(4)]
    smax = torch.zeros(1, **kw)
    L.call("nv_hydro_stage", *[t.data_ptr() for t in U], xi.data_ptr(), *[t.data_ptr() for t in U],
           *[t.data_ptr() for t in out], *shape, None, 0, f(0), None, f(0.9), f(0.04), f(1), f(1e-3), f(1 / 512),
           f(0), f(1), f(2e-3), f(1 / 144), f(1), f(6), 1, 0, smax.data_ptr(), stream(env))
    for c in range(4):
        assert bool((out[c] == U[c]).all()), f"field {c}"
    del U, out, xi
    torch.cuda.empty_cache()


def test_mass_change_equals_boundary_flux_512(env):
    """One RK stage of the 512^3 wall: the FP64 change of total mass equals -dt dx^2 times the net outward flux
    through the 6 boundary planes (clamped ghosts, nebula3d.py:23-25), to FP32 rounding of the cell updates."""
    import oracle_numpy as ON
    from nebulaviewer_amd import scenarios as S
    say("[property] mass telescoping at 512^3 (wall)")
    torch, dev, L = env
    setup = S.make_setup("wall", 512)
    shape = setup.shape
    dx = setup.dx
    rng = np.random.default_rng(21)
    rho = setup.rho
    U = [rho] + [(rho * rng.normal(0, 0.3, shape)).astype(F32) for _ in range(3)]
    xi = (rng.uniform(size=shape) < 0.5).astype(F32)
    dt = 2e-4
    cn2 = (1 / 12) ** 2
    gU = [T(env, u) for u in U]
    gxi = T(env, xi)
    out = [torch.empty(shape, dtype=torch.float32, device=dev) for _ in range(4)]
    L.call("nv_hydro_stage", *[t.data_ptr() for t in gU], gxi.data_ptr(), *[t.data_ptr() for t in gU],
           *[t.data_ptr() for t in out], *shape, None, 0, f(0), None, f(0.9), f(0.04), f(1), f(dt), f(dx), f(0),
           f(1), f(2e-3), f(cn2), f(1), f(6), 1, 0, None, stream(env))
    dM = float(out[0].double().sum() - gU[0].double().sum())
    # boundary faces in float64 from the float32 state; floor clamp never acts (rho >= 2e-3 by construction,
    # and the reconstruction floor is part of the face flux itself)
    U64 = [u.astype(np.float64) for u in U]
    r = np.maximum(U64[0], 2e-3)
    q = [r, U64[1] / r, U64[2] / r, U64[3] / r]
    c2 = cn2 + xi.astype(np.float64) * (1 - cn2)
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
    expect = -dt / dx * net
    # FP32 rounding of each cell's update: ~1 ulp of rho per cell, random-walk accumulated
    tol = 6 * np.sqrt(np.sum((np.spacing(U[0]).astype(np.float64)) ** 2)) + 1e-9 * abs(expect)
    assert abs(dM - expect) <= tol, (dM, expect, tol)
    assert abs(expect) > 10 * tol   # the boundary flux is resolved, not lost in rounding
    # the same budget from the GPU's FP64 sums: the float boundary fluxes of the stage against the host's FP64
    # restatement (float vs double face arithmetic: 1e-6), and the mass change against them to FP32 rounding
    sums_in = torch.zeros(2, dtype=torch.float64, device=dev)
    L.call("nv_conservation_sums", *[t.data_ptr() for t in gU], gxi.data_ptr(), *shape, f(2e-3), f(cn2), f(1.0), 1,
           sums_in.data_ptr(), stream(env))
    sums_out = torch.zeros(2, dtype=torch.float64, device=dev)
    L.call("nv_conservation_sums", *[t.data_ptr() for t in out], gxi.data_ptr(), *shape, f(2e-3), f(cn2), f(1.0), 1,
           sums_out.data_ptr(), stream(env))
    m_in, b_in = sums_in.tolist()
    m_out = sums_out.tolist()[0]
    assert abs(b_in - net) <= 1e-6 * abs(net), (b_in, net)
    assert abs((m_out - m_in) - (-dt / dx * b_in)) <= tol
    assert abs(m_in - float(gU[0].double().sum())) <= 1e-9 * m_in


def test_rk2_step_mass_budget(env):
    """one Sim step (both RK stages) on the 64^3 sheet (Mach-5 flow through every face; its wind adds momentum
    only): M^{n+1} - M^n = -(dt/dx) (B(U^n) + B(U1)) / 2 from the GPU's FP64 sums, to FP32 rounding (the density
    floor does not act)"""
    from nebulaviewer_amd import scenarios as S
    from nebulaviewer_amd.sim import Sim
    torch, dev, L = env
    say("[property] RK2 mass budget, sheet 64^3")
    sim = Sim(fast=False, setup=S.make_setup("sheet", 64, 1), verbose=False, rays=(64, 128), fuv_grid=(16, 16, 32))
    sim.step(3)                                           # some flow through the faces

    def sums(U):
        o = torch.zeros(2, dtype=torch.float64, device=dev)
        L.call("nv_conservation_sums", *[t.data_ptr() for t in U], sim.xi.data_ptr(), *sim.shape,
               f(sim.P["floor"]), f(sim.cn2), f(sim.ci2), 1, o.data_ptr(), stream(env))
        return o.tolist()

    m0, b0 = sums(sim.U)
    rho0 = sim.rho.cpu().numpy()
    assert float(sim.rho.min()) > 10 * sim.P["floor"]
    sim.step(1)
    m1, b1 = sums([sim.stage[c] for c in range(4)])       # U1: the stage block after the step
    m2 = sums(sim.U)[0]
    dt = sim.dt
    expect = -0.5 * float(np.float32(dt)) / sim.dx * (b0 + b1)
    tol = 12 * np.sqrt(np.sum(np.spacing(rho0).astype(np.float64) ** 2))
    assert abs((m2 - m0) - expect) <= tol, (m2 - m0, expect, tol)
    assert abs(expect) > 10 * tol


@pytest.mark.parametrize("nu,nphi", [(512, 1024), (24, 40), (7, 9)])
def test_ray_dirs_table(env, nu, nphi):
    """nv_ray_dirs against the ray direction of nebula3d.py:252-257 restated in float32 numpy with the oracle's
    shared sine and cosine, bit for bit (the batched casts read this table instead of computing it)"""
    from oracle.oracle import load
    say(f"[ray dirs] {nu}x{nphi}")
    torch, dev, L = env
    olib = load()
    tab = torch.empty(4 * nu * nphi, dtype=torch.float32, device=dev)
    L.call("nv_ray_dirs", nu, nphi, tab.data_ptr(), stream(env))
    torch.cuda.synchronize()
    got = N(tab).reshape(nu, nphi, 4)
    iu = np.arange(nu, dtype=F32)
    u = F32(-1) + (iu + F32(0.5)) * F32(2) / F32(nu)
    s = np.sqrt(np.maximum(F32(1) - u * u, F32(0)))
    phi = (np.arange(nphi, dtype=F32) + F32(0.5)) * F32(2) * F32(math.pi) / F32(nphi)
    sn = np.array([olib.or_sinf_det(float(v)) for v in phi], F32)
    cs = np.array([olib.or_cosf_det(float(v)) for v in phi], F32)
    np.testing.assert_array_equal(got[..., 0], s[:, None] * cs[None, :])
    np.testing.assert_array_equal(got[..., 1], s[:, None] * sn[None, :])
    np.testing.assert_array_equal(got[..., 2], np.broadcast_to(u[:, None], (nu, nphi)))
    assert not got[..., 3].any()
