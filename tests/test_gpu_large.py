"""Full-size properties of the 1024^3 grid (1.07 G cells; fields of 4.3 GB, so every byte offset past the 512th
x-plane exceeds 2^31): one FP32 RK stage (hydro_stage_hll, sim.py:423-498, launched as in sim.py:1017-1025)
keeps a uniform state bitwise steady, changes the FP64 total mass by exactly the boundary-flux term to FP32
rounding, and gives the CPU oracle's bits on the grid's first and last x-planes. The state is made on the device
(a seeded lognormal density with random velocities), not on the host."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 1024
f = ctypes.c_float
CN2 = (1 / 12) ** 2


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def env():
    import torch
    from nebulaviewer_amd import _lib
    return torch, torch.device("cuda:0"), _lib


def stage(env, U, xi, out, dt, smax=None):
    torch, dev, L = env
    strm = torch.cuda.current_stream(dev).cuda_stream
    L.call("nv_hydro_stage", *[t.data_ptr() for t in U], xi.data_ptr(), *[t.data_ptr() for t in U],
           *[t.data_ptr() for t in out], N, N, N, None, 0, f(0), None, f(0.9), f(0.04), f(1), f(dt), f(1 / N), f(0),
           f(1), f(2e-3), f(CN2), f(1), f(6), 1, 0, None if smax is None else smax.data_ptr(), strm)


def test_uniform_state_is_steady_1024(env, okern32):
    torch, dev, L = env
    say("[large] uniform flow at 1024^3: one stage leaves it bitwise unchanged")
    kw = dict(dtype=torch.float32, device=dev)
    shape = (N, N, N)
    U = [torch.full(shape, v, **kw) for v in (3.0, 0.6, -0.3, 0.15)]
    xi = torch.full(shape, 0.25, **kw)
    out = [torch.empty(shape, **kw) for _ in range(4)]
    smax = torch.zeros(1, **kw)
    stage(env, U, xi, out, 1e-3, smax)
    for c in range(4):
        assert bool((out[c] == U[c]).all()), f"field {c}"
    # the fused CFL maximum: |u| + |v| + |w| + c of the uniform flow, as the oracle's max_speed gives it
    tiny = [np.full((2, 2, 2), v, np.float32) for v in (3.0, 0.6, -0.3, 0.15, 0.25)]
    assert float(smax.item()) == okern32.max_speed(*tiny, CN2, 1.0)
    del U, out, xi
    torch.cuda.empty_cache()


def test_mass_budget_and_edge_planes_1024(env, okern32):
    torch, dev, L = env
    say("[large] 1024^3 random state on the device")
    shape = (N, N, N)
    kw = dict(dtype=torch.float32, device=dev)
    g = torch.Generator(device=dev)
    g.manual_seed(3)
    rho = torch.empty(shape, **kw)
    rho.normal_(0.0, 0.4, generator=g).exp_()
    U = [rho]
    for _ in range(3):
        m = torch.empty(shape, **kw)
        m.normal_(0.0, 0.3, generator=g)
        U.append(m)
    # an outflow of 0.2 through both x faces on top of the random velocities, so that the boundary term is large
    U[1].add_(torch.where(torch.arange(N, device=dev) < N // 2, -0.2, 0.2).to(torch.float32).view(N, 1, 1))
    for m in U[1:]:
        m.mul_(rho)
    xi = (torch.rand(shape, generator=g, **kw) < 0.5).float()
    out = [torch.empty(shape, **kw) for _ in range(4)]
    dt = 1e-4
    strm = torch.cuda.current_stream(dev).cuda_stream

    def sums(F):
        o = torch.zeros(2, dtype=torch.float64, device=dev)
        L.call("nv_conservation_sums", *[t.data_ptr() for t in F], xi.data_ptr(), N, N, N, f(2e-3), f(CN2), f(1.0), 1,
               o.data_ptr(), strm)
        return o.tolist()

    say("[large] one FP32 stage and the FP64 sums")
    stage(env, U, xi, out, dt)
    m_in, b_in = sums(U)
    m_out = sums(out)[0]
    # FP32 rounding of each cell's update (half an ulp of rho at most), random-walk accumulated
    tol = 12.0 * float(torch.sqrt(torch.sum((rho.double() * 2.0 ** -24) ** 2)))
    expect = -dt * N * b_in
    say(f"[large] dM = {m_out - m_in:.6e}, boundary term {expect:.6e}, tolerance {tol:.3e}")
    assert abs((m_out - m_in) - expect) <= tol, (m_out - m_in, expect, tol)
    assert abs(expect) > 10 * tol            # the boundary flux is resolved, not lost in rounding
    assert float(torch.min(out[0])) >= 2e-3

    # the first and last six x-planes against the oracle on 8-plane slabs: a slab's clamped ghosts at its cut face
    # reach 2 planes in, so planes 0-5 of [0, 8) and 1018-1023 of [1016, 1024) are the full grid's own
    z = np.zeros((1, 4), np.float32)
    for lo, keep in ((0, slice(0, 6)), (N - 8, slice(2, 8))):
        say(f"[large] oracle on x-planes {lo}..{lo + 7}")
        Us = [np.ascontiguousarray(t[lo:lo + 8].cpu().numpy()) for t in U]
        xs = np.ascontiguousarray(xi[lo:lo + 8].cpu().numpy())
        ref = [np.zeros_like(Us[0]) for _ in range(4)]
        okern32.hydro_stage(Us, xs, Us, ref, Us[0].shape, z, 0, 0.0, z, 0.9, 0.04, 1.0, dt, 1 / N, 0.0, 1.0, 2e-3,
                            CN2, 1.0, 6.0, 1, 0)
        for c in range(4):
            got = out[c][lo:lo + 8].cpu().numpy()
            np.testing.assert_array_equal(got[keep], ref[c][keep], err_msg=f"field {c}, planes from {lo}")
    del U, out, xi, rho
    torch.cuda.empty_cache()
