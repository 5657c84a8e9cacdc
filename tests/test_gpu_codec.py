"""The consumers' kernels on the GPU against the oracle, bit for bit: the timeline codec (both dfront / g0
storages), the shock proxy, the render pack (shock on/off, fade on/off) and prefilter, and the bake frame's
stored forms against the reference's own numpy expressions (bake.py:94-99). Then the host classes built on them:
a Timeline capture/restore round trip and branch, and a baked frame set."""
import ctypes
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
F32 = np.float32


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def env():
    import torch
    from nebulaviewer_amd import _lib
    from oracle.oracle import Codec
    _lib.load()
    return torch, _lib, Codec()


def G(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def N(t):
    return t.cpu().numpy()


def fields(shape, seed):
    rng = np.random.default_rng(seed)
    rho = np.exp(rng.normal(0, 2, shape)).astype(F32)
    rho.flat[:3] = [7e4, 1e-5, 65000]
    m = [(rho * rng.normal(0, 2, shape)).astype(F32) for _ in range(3)]
    m[0].flat[:2] = [9e4, -9e4]
    xi = rng.uniform(-0.2, 1.2, shape).astype(F32)
    df = rng.normal(0, 3, shape).astype(F32)
    g0 = np.exp(rng.normal(0, 3, shape)).astype(F32)
    return rho, m, xi, df, g0


@pytest.mark.parametrize("half", [True, False])
def test_timeline_codec_bitwise(env, half):
    torch, L, C = env
    shape = (19, 23, 37)
    say(f"[codec] timeline half={half}")
    rho, m, xi, df, g0 = fields(shape, 3)
    if half:
        df, g0 = df.astype(np.float16), g0.astype(np.float16)
    ref = C.timeline_pack(rho, *m, xi, df.astype(F32), g0.astype(F32))
    gin = [G(torch, a) for a in (rho, *m, xi, df, g0)]
    out = [torch.empty(shape, dtype=torch.float16, device="cuda:0") for _ in range(4)] + \
          [torch.empty(shape, dtype=torch.uint8, device="cuda:0")] + \
          [torch.empty(shape, dtype=torch.float16, device="cuda:0") for _ in range(2)]
    n = rho.size
    L.call("nv_timeline_pack", *[t.data_ptr() for t in gin], int(half), *[t.data_ptr() for t in out], n, None)
    torch.cuda.synchronize()
    for k in range(7):
        got = N(out[k]) if k == 4 else N(out[k]).view(np.uint16)
        np.testing.assert_array_equal(got, ref[k], err_msg=f"array {k}")
    back_ref = C.timeline_unpack(ref, 2e-3)
    dt = torch.float16 if half else torch.float32
    back = [torch.empty(shape, dtype=torch.float32, device="cuda:0") for _ in range(5)] + \
           [torch.empty(shape, dtype=dt, device="cuda:0") for _ in range(2)]
    L.call("nv_timeline_unpack", *[t.data_ptr() for t in out], ctypes.c_float(2e-3), *[t.data_ptr() for t in back],
           int(half), n, None)
    torch.cuda.synchronize()
    for k in range(7):
        np.testing.assert_array_equal(N(back[k]).astype(F32), back_ref[k], err_msg=f"field {k}")


@pytest.mark.parametrize("shape,dx", [((16, 20, 24), 1 / 24), ((9, 5, 70), 0.013), ((1, 1, 5), 0.2)])
def test_shock_bitwise(env, shape, dx):
    torch, L, C = env
    say(f"[codec] shock {shape}")
    rho, m, _, _, _ = fields(shape, 4)
    rho = np.abs(rho) + F32(0.01)
    ref = C.shock_field(rho, *m, dx)
    gin = [G(torch, a) for a in (rho, *m)]
    out = torch.empty(shape, dtype=torch.float32, device="cuda:0")
    L.call("nv_shock_field", *[t.data_ptr() for t in gin], *shape, ctypes.c_float(dx), out.data_ptr(), None)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(N(out), ref)


@pytest.mark.parametrize("half,shock_on,edge", [(True, 1, 0.1), (False, 0, 0.0), (True, 0, 0.37)])
def test_pack_volume_and_smooth_bitwise(env, half, shock_on, edge):
    torch, L, C = env
    shape = (14, 18, 22)
    say(f"[codec] pack half={half} shock={shock_on} edge={edge}")
    rho, m, xi, df, g0 = fields(shape, 5)
    rho = np.abs(rho) + F32(0.01)
    df = (df * F32(0.05)).astype(F32)                    # around the front, where the line weights vary
    if half:
        df, g0 = df.astype(np.float16), np.minimum(g0, 6e4).astype(np.float16)
    dx = 1.0 / 22
    pp = L.NvPackParams()
    vals = dict(c_ha=(1.0, 0.3, 0.1), c_o3=(0.0, 0.9, 0.7), c_s2=(0.8, 0.1, 0.0), c_dg=(0.05, 0.04, 0.03),
                c_sh=(0.2, 0.2, 0.5) if shock_on else (0.0, 0.0, 0.0))
    for k, v in vals.items():
        getattr(pp, k)[:] = v
    pp.gs, pp.ambient, pp.edge = 7.0, 0.03, edge
    pp.inv_L[:] = [1 / (shape[0] * dx), 1 / (shape[1] * dx), 1 / (shape[2] * dx)]
    pp.dx = dx
    flat = np.array(list(vals["c_ha"]) + list(vals["c_o3"]) + list(vals["c_s2"]) + list(vals["c_dg"]) +
                    list(vals["c_sh"]) + [7.0, 0.03, edge] + list(pp.inv_L) + [dx], F32)
    ref = C.pack_volume(rho, *m, xi, df.astype(F32), g0.astype(F32), flat, shock_on)
    gin = [G(torch, a) for a in (rho, *m, xi, df, g0)]
    out = torch.empty(shape + (4,), dtype=torch.float16, device="cuda:0")
    L.call("nv_pack_volume", *[t.data_ptr() for t in gin], int(half), ctypes.byref(pp), shock_on, out.data_ptr(),
           *shape, None)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(N(out).view(np.uint16), ref)
    for axis in range(3):
        sm = torch.empty_like(out)
        L.call("nv_smooth_axis", out.data_ptr(), sm.data_ptr(), axis, *shape, None)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(N(sm).view(np.uint16), C.smooth_axis(ref, axis), err_msg=f"axis {axis}")


@pytest.mark.parametrize("half", [True, False])
def test_frame_pack_matches_bake_numpy(env, half):
    """the stored forms are the reference bake's numpy expressions on the same fields (bake.py:94-99)"""
    torch, L, C = env
    shape = (12, 13, 14)
    say(f"[codec] frame half={half}")
    rho, _, xi, df, g0 = fields(shape, 6)
    sh = np.exp(np.random.default_rng(7).normal(3, 4, shape)).astype(F32)
    xi.flat[:4] = [0.5 / 255, 1.5 / 255, 2.5 / 255, 1.0]       # exact-looking halves: numpy rounds to even
    if half:
        df, g0 = df.astype(np.float16), g0.astype(np.float16)
    gin = [G(torch, a) for a in (rho, xi, df, g0, sh)]
    h = lambda: torch.empty(shape, dtype=torch.float16, device="cuda:0")
    o_rho, o_df, o_g0, o_sh = h(), h(), h(), h()
    o_xi = torch.empty(shape, dtype=torch.uint8, device="cuda:0")
    L.call("nv_frame_pack", *[t.data_ptr() for t in gin], int(half), o_rho.data_ptr(), o_xi.data_ptr(),
           o_df.data_ptr(), o_g0.data_ptr(), o_sh.data_ptr(), rho.size, None)
    torch.cuda.synchronize()
    with np.errstate(over="ignore"):
        want = dict(g0=np.minimum(g0, 6e4).astype(np.float16), rho=rho.astype(np.float16),
                    xi=np.round(np.clip(xi, 0, 1) * 255).astype(np.uint8), df=np.clip(df, -4, 4).astype(np.float16),
                    sh=np.minimum(sh, 6e4).astype(np.float16))
    got = dict(g0=N(o_g0), rho=N(o_rho), xi=N(o_xi), df=N(o_df), sh=N(o_sh))
    for k in want:
        a, b = got[k], want[k]
        if a.dtype == np.float16:
            a, b = a.view(np.uint16), b.view(np.uint16)
        np.testing.assert_array_equal(a, b, err_msg=k)


def test_timeline_capture_restore_and_branch(env):
    """restore gives back the packed state exactly (the codec's fixed point), with the clock and star tables;
    stepping on from a restore is the same as stepping on from an identical restore"""
    torch, L, C = env
    from nebulaviewer_amd import scenarios as S
    from nebulaviewer_amd.sim import Sim
    from nebulaviewer_amd.timeline import Timeline
    say("[timeline] capture / restore / branch on a torus 32^3")
    sim = Sim(fast=False, setup=S.make_setup("torus", 32), verbose=False, rays=(32, 64), fuv_grid=(16, 16, 32))
    tl = Timeline(budget_gb=1.0)
    tl.clear(sim)
    tl.dtau = 1e-9                    # a snapshot after every step
    for _ in range(6):
        sim.step(1)
        assert tl.maybe_capture(sim)
    assert len(tl.snaps) == 6 and tl.taus() == sorted(tl.taus())
    t3 = tl.snaps[3].tau
    assert tl.nearest(t3 + 1e-12) == 3 and tl.index_at_or_before(t3) == 3
    # restore snapshot 3, pack again: identical stored arrays
    assert tl.restore(sim, 3)
    assert sim.t == t3
    from nebulaviewer_amd.timeline import pack_state
    again = pack_state(sim)
    for a, b in zip(again, tl.snaps[3].arrays):
        assert torch.equal(a, b)
    sim.step(2)
    r1 = sim.rho.clone()
    tl.restore(sim, 3)
    sim.step(2)
    assert torch.equal(sim.rho, r1)
    tl.truncate_after(t3)
    assert len(tl.snaps) == 4
    tl.set_budget(2.5 * 32 ** 3 * 13 / 1e9)        # room for two snapshots
    assert len(tl.snaps) == 2 and tl.thinned == 2         # 4 -> [0, 2, 3] -> [0, 3]


def test_bake_frames(env, tmp_path):
    torch, L, C = env
    from nebulaviewer_amd import scenarios as S
    from nebulaviewer_amd.sim import Sim
    from nebulaviewer_amd.bake import bake
    say("[bake] 3 frames of a torus 24^3")
    sim = Sim(fast=False, setup=S.make_setup("torus", 24), verbose=False, rays=(32, 64), fuv_grid=(16, 16, 32))
    t_end = 5 * 0.3 / 24 / 12         # a handful of CFL steps
    meta = bake(sim, str(tmp_path), t_end, 3)
    files = sorted(os.listdir(tmp_path / "frames"))
    assert len(files) == 15
    for f in range(3):
        for name, dt in (("rho", np.float16), ("xi", np.uint8), ("df", np.float16), ("g0", np.float16),
                         ("sh", np.float16)):
            a = np.load(tmp_path / "frames" / f"f{f:04d}_{name}.npy")
            assert a.dtype == dt and a.shape == (24, 24, 24)
    m = json.load(open(tmp_path / "meta.json"))
    assert m["shape"] == [24, 24, 24] and len(m["times"]) == 3
    assert abs(m["times"][-1] - t_end) < 1e-6 and m["times"] == sorted(m["times"])
    assert meta["steps"] == sim.it
