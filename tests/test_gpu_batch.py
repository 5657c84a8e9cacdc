"""The batched step (SimBatch: nv_hydro_stage_batch, nv_cast_rays_batch, nv_ionize_apply_batch, nv_cfl_dt)
against each member stepped alone by Sim: bit-identical fields and clocks, on uneven grids, with the device
clock, with HIP-graph replay, and with the host clock (a member whose luminosity ramps)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RAYS = (32, 64)


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    return torch


def setups(params=None):
    from nebulaviewer_amd import scenarios as S
    out = [S.make_setup("torus", 24, 0), S.make_setup("sheet", 32, 1), S.make_setup("torus", 40, 2),
           S.make_setup("sheet", 20, 5, dict(wind=0.9))]
    if params:
        for st, p in zip(out, params):
            if p:
                st.P.update(p)
    return out


def same(a, b, what):
    a, b = a.float().cpu().numpy(), b.float().cpu().numpy()
    assert np.array_equal(a, b), f"{what}: {np.abs(a - b).max()}"


def check_members(batch, steps, precision, params=None):
    from nebulaviewer_amd.sim import Sim
    for k, m in enumerate(batch.members):
        ref = Sim(setup=setups(params)[k], verbose=False, precision=precision, rays=RAYS)
        ref.step(steps)
        for name in ("rho", "mx", "my", "mz", "xi", "dfront"):
            same(getattr(m, name), getattr(ref, name), f"member {k} {name}")
        assert m.t == ref.t and m.it == ref.it, (k, m.t, ref.t)


@pytest.mark.parametrize("precision", ["fp32", "fp32-fast"])
def test_batch_device_clock_matches_single(torch_dev, precision):
    from nebulaviewer_amd.batch import SimBatch
    say(f"[batch] device clock, {precision}")
    b = SimBatch(setups(), precision=precision, rays=RAYS)
    assert b.device_clock_ok()
    b.step(5)
    check_members(b, 5, precision)


def test_batch_graph_replay_matches_single(torch_dev):
    """every replay of the two captured steps (with and without ionization) against an eager batch stepped
    beside it, with eager work and a synchronization between replays; then against the single Sims"""
    from nebulaviewer_amd.batch import SimBatch
    torch = torch_dev
    say("[batch] HIP graph replay")
    g = SimBatch(setups(), rays=RAYS, graph=True)
    e = SimBatch(setups(), rays=RAYS, graph=False)
    for k in range(7):
        g.step(1)
        e.step(1)
        torch.cuda.synchronize()
        for s, (mg, me) in enumerate(zip(g.members, e.members)):
            for name in ("rho", "mx", "xi", "dfront"):
                same(getattr(mg, name), getattr(me, name), f"step {k} member {s} {name}")
            assert mg.t == me.t
    assert set(g._graphs) == {True, False}
    check_members(g, 7, "fp32")


def test_batch_host_clock_matches_single(torch_dev):
    from nebulaviewer_amd.batch import SimBatch
    say("[batch] host clock (luminosity ramp)")
    params = [None, dict(q_ramp=0.05), None, None]
    b = SimBatch(setups(params), rays=RAYS)
    assert not b.device_clock_ok()
    b.step(5)
    check_members(b, 5, "fp32", params)


def test_batch_fp16_velocity_matches_single(torch_dev):
    from nebulaviewer_amd.batch import SimBatch
    say("[batch] fp16-vel")
    b = SimBatch(setups(), precision="fp16-vel", rays=RAYS)
    b.step(4)
    check_members(b, 4, "fp16-vel")


def test_autotuned_tiles_change_no_bit(torch_dev):
    """The measured tile choice (Sim / SimBatch tile="tune") against every fixed tile: identical fields and clocks;
    the tuner timed all three shapes and kept the fastest."""
    from nebulaviewer_amd import autotune
    from nebulaviewer_amd.batch import SimBatch
    from nebulaviewer_amd.sim import Sim
    from nebulaviewer_amd import scenarios as S
    say("[autotune] Sim 40^3 torus and a batch, tuned vs fixed tiles")
    runs = {}
    for tile in ("tune", 0, 1, 2, 3):
        sim = Sim(setup=S.make_setup("torus", 40, 0), verbose=False, rays=RAYS, tile=tile)
        sim.step(4)
        runs[tile] = sim
    for tile in (0, 1, 2, 3):
        for name in ("rho", "mx", "my", "mz", "xi"):
            same(getattr(runs["tune"], name), getattr(runs[tile], name), f"tile {tile} {name}")
        assert runs["tune"].t == runs[tile].t
    assert runs["tune"].tile_shape in (1, 2, 3)
    rep = autotune.report()
    key = [k for k in rep if "'sim', (40, 40, 40)" in k]
    assert key and set(rep[key[0]]["ms"]) == {"64x8", "32x16", "64x16"}
    b = SimBatch(setups(), rays=RAYS)
    b.step(3)
    assert b.tile in (1, 2, 3)
    check_members(b, 3, "fp32")


def test_batch_mixed_brick_jobs_match_single(torch_dev):
    """Ray jobs with and without density bricks in one batched cast (member 1 samples rho directly; the sheets'
    four stars share their member's brick array, built once): every member equals its single Sim."""
    from nebulaviewer_amd.batch import SimBatch
    say("[batch] mixed brick / direct ray jobs")
    b = SimBatch(setups(), rays=RAYS)
    b.ray_bricks = True
    b.members[1].ray_bricks = False
    b.step(3)
    assert b.ray_bricks_used
    check_members(b, 3, "fp32")


@pytest.mark.parametrize("precision", ["fp32", "fp16-vel"])
def test_batch_queue_dispatch_matches_single(torch_dev, precision):
    """The persistent tile-queue launch of both stages (nv_hydro_stage_batch_queue) against each member stepped
    alone, over several steps and two kinds of captured graph (the queue's counters reset after every launch)."""
    from nebulaviewer_amd.batch import SimBatch
    say(f"[batch] persistent tile-queue dispatch, {precision}")
    b = SimBatch(setups(), precision=precision, rays=RAYS, dispatch="queue", graph=True)
    assert b.dispatch == {1: True, 2: True}
    b.step(6)
    check_members(b, 6, precision)


def test_batch_dispatch_is_tuned(torch_dev):
    from nebulaviewer_amd.batch import SimBatch
    from nebulaviewer_amd import autotune
    say("[batch] dispatch tuning")
    b = SimBatch(setups(), rays=RAYS)
    b.step(2)
    assert set(b.dispatch) == {1, 2}
    keys = [k for k in autotune.report() if "batch_dispatch" in k]
    assert len(keys) >= 2
    check_members(b, 2, "fp32")
