"""The pre-turbulence kernels against the oracle (nv_drive, nv_add_torus bitwise; nv_mach_sums to FP64
summation order; nv_every2 against numpy slicing), and a Sim run through the phase change - stirring, the
Mach estimate, the torus, then ionized steps with stirring kept on - bitwise against OracleSim."""
import ctypes

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


@pytest.mark.parametrize("dxs", [1.0 / 24.0, 1.0 / 32.0])
def test_drive_bitwise(env, dxs):
    torch, L, C = env
    say(f"[pre] drive dxs={dxs}")
    n = 40
    shape = (n, 36, 44)
    dx = 1.0 / 44
    sshape = tuple(int(np.ceil(s * dx / dxs)) + 2 for s in shape)
    rng = np.random.default_rng(8)
    rho = np.exp(rng.normal(0, 1, shape)).astype(F32)
    m = [rng.normal(0, 1, shape).astype(F32) for _ in range(3)]
    F = [rng.normal(0, 1, sshape).astype(F32) for _ in range(3)]
    Gs = [rng.normal(0, 1, sshape).astype(F32) for _ in range(3)]
    gm = [G(torch, a) for a in m]
    grho = G(torch, rho)
    gF, gG = [G(torch, a) for a in F], [G(torch, a) for a in Gs]
    C.drive(rho, m, F, Gs, 0.37, 7.0, 0.003, dx, dxs)
    L.call("nv_drive", grho.data_ptr(), *[t.data_ptr() for t in gm], *shape, *[t.data_ptr() for t in gF],
           *[t.data_ptr() for t in gG], *sshape, ctypes.c_float(0.37), ctypes.c_float(7.0), ctypes.c_float(0.003),
           ctypes.c_float(dx), ctypes.c_float(dxs), None)
    torch.cuda.synchronize()
    for c in range(3):
        np.testing.assert_array_equal(gm[c].cpu().numpy(), m[c])


def test_torus_mach_every2(env):
    torch, L, C = env
    say("[pre] add_torus / mach_sums / every2")
    shape = (33, 20, 27)
    rng = np.random.default_rng(9)
    rho = np.exp(rng.normal(0, 1, shape)).astype(F32)
    m = [rng.normal(0, 1, shape).astype(F32) for _ in range(3)]
    tor = np.exp(rng.normal(0, 1, shape)).astype(F32)
    g = [G(torch, a) for a in (rho, *m)]
    gt = G(torch, tor)
    sums = torch.zeros(2, dtype=torch.float64, device="cuda:0")
    L.call("nv_mach_sums", *[t.data_ptr() for t in g], rho.size, sums.data_ptr(), None)
    ref = C.mach_sums(rho, m)
    np.testing.assert_allclose(sums.cpu().numpy(), ref, rtol=1e-12)
    C.add_torus(rho, m, tor)
    L.call("nv_add_torus", *[t.data_ptr() for t in g], gt.data_ptr(), rho.size, None)
    torch.cuda.synchronize()
    for a, b in zip(g, (rho, *m)):
        np.testing.assert_array_equal(a.cpu().numpy(), b)
    for half in (False, True):
        src = G(torch, rho.astype(np.float16) if half else rho)
        out = torch.empty(tuple((s + 1) // 2 for s in shape), dtype=torch.float32, device="cuda:0")
        L.call("nv_every2", src.data_ptr(), int(half), out.data_ptr(), *shape, None)
        torch.cuda.synchronize()
        want = (rho.astype(np.float16) if half else rho)[::2, ::2, ::2].astype(F32)
        np.testing.assert_array_equal(out.cpu().numpy(), want)


def test_sim_through_pre_turbulence_matches_oracle(env):
    from nebulaviewer_amd import scenarios as S
    from nebulaviewer_amd.sim import Sim
    from oracle.oracle import OracleSim
    say("[pre] Sim torus 24^3 with a turbulent phase, 9 steps, vs OracleSim")
    params = dict(pre_t=0.16, drive_main=0.5)   # 5 pre-turbulence steps, then 4 ionized ones
    st = S.make_setup("torus", 24, params=params)
    sim = Sim(fast=False, setup=st, verbose=False, rays=(32, 64), fuv_grid=(16, 16, 32))
    ref = OracleSim(S.make_setup("torus", 24, params=params), S.constants(st), (32, 64), (16, 16, 32), np.float32,
                    threads=8)
    sim.step(9)
    ref.step(9)
    assert ref.phase == "main" and ref.ip >= 2 and ref.it >= 2, (ref.ip, ref.it)
    assert sim.phase == ref.phase and sim.ip == ref.ip and sim.it == ref.it
    assert sim.tp == ref.tp and sim.t == ref.t
    assert abs(sim.mach - ref.mach) <= 1e-9 * ref.mach
    for name in ("rho", "mx", "my", "mz", "xi"):
        np.testing.assert_array_equal(getattr(sim, name).cpu().numpy(), getattr(ref, name), err_msg=name)
