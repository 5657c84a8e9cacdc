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
        ref = Sim(fast=False, setup=setups(params)[k], verbose=False, precision=precision, rays=RAYS)
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


@pytest.mark.parametrize("tile", [2 | (1 << 15 << 8), 1 | (5 << 8)])
def test_batch_march_depth_matches_single(torch_dev, tile):
    """batch stage plans whose workgroups march whole member columns, or 5-plane x chunks (NV_TILE_DEPTH), with
    both dispatch forms: every member equals its single Sim"""
    from nebulaviewer_amd.batch import SimBatch
    for dispatch in ("hardware", "queue"):
        say(f"[batch] tile code {tile:#x}, {dispatch}")
        b = SimBatch(setups(), precision="fp32", rays=RAYS, tile=tile, dispatch=dispatch)
        b.step(4)
        check_members(b, 4, "fp32")


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
    fixed = (0, 1, 2, 3, 3 | (autotune.WHOLE_COLUMN << 8), 1 | (7 << 8))   # shapes, and shapes with march depths
    for tile in ("tune",) + fixed:
        sim = Sim(fast=False, setup=S.make_setup("torus", 40, 0), verbose=False, rays=RAYS, tile=tile)
        sim.step(4)
        runs[tile] = sim
    for tile in fixed:
        for name in ("rho", "mx", "my", "mz", "xi"):
            same(getattr(runs["tune"], name), getattr(runs[tile], name), f"tile {tile} {name}")
        assert runs["tune"].t == runs[tile].t
    assert len(runs["tune"].tile_shape) == 2 and all(t & 0xFF in (1, 2, 3) for t in runs["tune"].tile_shape)
    rep = autotune.report()
    keys = [k for k in rep if "'sim', (40, 40, 40), 'fp32', " in k]
    names = {autotune.tile_name(c) for c in autotune.SIM_CANDIDATES}
    assert names == {"64x8", "32x16", "64x16", "64x8/whole", "32x16/whole", "64x16/whole"}
    assert len(keys) == 2 and all(set(rep[k]["ms"]) == names for k in keys)
    b = SimBatch(setups(), rays=RAYS)
    b.step(3)
    assert set(b.tile) == {1, 2} and all(t & 0xFF in (1, 2, 3) for t in b.tile.values())
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


def test_batch_queue_one_workgroup_per_cu_matches_single(torch_dev):
    """The persistent launch at one resident workgroup per CU (nv_hydro_stage_batch_queue_ex: 8 waves per CU on
    these 512-thread tiles, half the full residency) gives the members' single-grid steps bit for bit."""
    from nebulaviewer_amd.batch import SimBatch
    say("[batch] persistent tile queue at one workgroup per CU")
    b = SimBatch(setups(), rays=RAYS, dispatch="queue", queue_wgs=1, graph=True)
    assert b.dispatch == {1: True, 2: True} and b.queue_wgs == {1: 1, 2: 1}
    b.step(4)
    check_members(b, 4, "fp32")


def test_batch_dispatch_is_tuned(torch_dev):
    from nebulaviewer_amd.batch import SimBatch
    from nebulaviewer_amd import autotune
    say("[batch] dispatch tuning")
    b = SimBatch(setups(), rays=RAYS)
    b.step(2)
    assert set(b.dispatch) == {1, 2} and set(b.queue_wgs) == {1, 2}
    keys = [k for k in autotune.report() if "batch_dispatch" in k]
    assert len(keys) >= 2
    for k in keys:   # hardware dispatch, the queue at full residency and at one workgroup per CU, all measured
        assert set(autotune.report()[k]["ms"]) == {"hardware", "queue", "queue-1wg/CU"}
    check_members(b, 2, "fp32")


# ------------------------------------------------------------------ the uneven batch against the oracle
# Members of scenarios.batch_setups at their real sizes: every size 48-160, torus and sheet, wind > 0 on all.
ORACLE_MEMBERS = (7, 12, 1, 16, 2, 15, 24, 43)


def test_batch_members_match_oracle(torch_dev):
    """The uneven batch (persistent tile-queue stages, HIP-graph replay, the real 512x1024 ray grid) against the
    CPU oracle: each member, stepped 5 times through three ionizations, equals OracleSim on its own setup bit
    for bit (sim.py:1099-1114 per member)."""
    import os
    from nebulaviewer_amd import scenarios as S
    from nebulaviewer_amd.batch import SimBatch
    from oracle.oracle import OracleSim
    steps = 5
    say(f"[batch-oracle] members {ORACLE_MEMBERS}: building the batch")
    b = SimBatch([S.batch_setup(i) for i in ORACLE_MEMBERS], dispatch="queue", graph=True)
    assert b.dispatch == {1: True, 2: True} and b.device_clock_ok()
    assert sorted({m.shape[0] for m in b.members}) == [48, 64, 96, 128, 160]
    assert {m.scenario for m in b.members} == {"torus", "sheet"} and all(m.P["wind"] > 0 for m in b.members)
    b.step(steps)
    torch_dev.cuda.synchronize()
    threads = min(16, os.cpu_count() or 1)
    for i, m in zip(ORACLE_MEMBERS, b.members):
        st = S.batch_setup(i)
        say(f"[batch-oracle] member {i}: {st.name} {st.shape[0]}^3, {len(st.stars)} stars, wind "
            f"{st.P['wind']:.2f}: oracle {steps} steps")
        ref = OracleSim(st, S.constants(st), S.ray_grid(st.shape[2]), S.fuv_grid(st.shape[2]), np.float32,
                        threads=threads)
        ref.step(steps)
        for name in ("rho", "mx", "my", "mz", "xi"):
            a = getattr(m, name).cpu().numpy()
            assert np.array_equal(a, getattr(ref, name)), \
                f"member {i} {name}: max |d| {np.abs(a - getattr(ref, name)).max()}"
        assert np.array_equal(m.dfront.cpu().numpy().view(np.uint16), ref.dfront.view(np.uint16)), f"member {i} dfront"
        assert m.t == ref.t and m.it == ref.it, (i, m.t, ref.t)


def test_full_batch_mass_budget(torch_dev):
    """All 64 members stepped by graph replay; on a step without ionization each member's FP64 mass change equals
    its two stages' boundary-flux term (nv_conservation_sums) to the rounding of the cells' float updates."""
    import ctypes
    import torch
    from nebulaviewer_amd import _lib
    from nebulaviewer_amd import scenarios as S
    from nebulaviewer_amd.batch import SimBatch
    f = ctypes.c_float
    say("[batch-mass] 64 members: initial conditions")
    b = SimBatch(S.batch_setups(64), graph=True)
    b.step(1)                                  # the ionization step
    assert b.it % b.ion_every != 0

    def sums(m, U):
        o = torch.zeros(2, dtype=torch.float64, device=m.device)
        _lib.call("nv_conservation_sums", *[t.data_ptr() for t in U], m.xi.data_ptr(), *m.shape, f(m.P["floor"]),
                  f(m.cn2), f(m.ci2), 1, o.data_ptr(), m._stream)
        return o.tolist()

    before = [sums(m, m.U) for m in b.members]
    bounds = [12.0 * float(torch.sqrt(torch.sum((m.rho.double() * 2.0 ** -23) ** 2))) for m in b.members]
    floor_hit = [float(m.rho.min()) <= 1.0001 * m.P["floor"] for m in b.members]
    say("[batch-mass] one step without ionization")
    b.step(1)
    torch.cuda.synchronize()
    checked = 0
    for s, m in enumerate(b.members):
        b1 = sums(m, [m.stage[c] for c in range(4)])[1]
        m2 = sums(m, m.U)[0]
        m0, b0 = before[s]
        flux = -0.5 * float(f(m.dt).value) / m.dx * (b0 + b1)
        if floor_hit[s]:
            continue
        checked += 1
        assert abs((m2 - m0) - flux) <= bounds[s], (s, m2 - m0, flux, bounds[s])
    assert checked >= 60


# ------------------------------------------------------------------ batch bookkeeping
def test_batch_refuses_unbatched_phases(torch_dev):
    """A member in its pre-turbulence phase, or stirred in the main phase, is refused (SimBatch steps the
    main-phase step only)."""
    from nebulaviewer_amd import scenarios as S
    from nebulaviewer_amd.batch import SimBatch
    say("[batch] refusal of pre-phase / stirred members")
    with pytest.raises(ValueError, match="phase"):
        SimBatch([S.make_setup("torus", 24, 0, dict(pre_t=0.05)), S.make_setup("sheet", 20, 1)], rays=RAYS)
    with pytest.raises(ValueError, match="drive_main"):
        SimBatch([S.make_setup("torus", 24, 0, dict(drive_main=0.5)), S.make_setup("sheet", 20, 1)], rays=RAYS)


def test_batch_host_clock_then_single_steps(torch_dev):
    """Host-clock batch steps followed by steps of one member on its own: the member's own CFL maximum is taken
    afresh (no stale cached value), so it still equals Sim stepped alone."""
    from nebulaviewer_amd.batch import SimBatch
    from nebulaviewer_amd.sim import Sim
    say("[batch] host-clock batch steps, then single-member steps")
    params = [dict(q_ramp=0.05), None, None, None]
    b = SimBatch(setups(params), rays=RAYS)
    assert not b.device_clock_ok()
    b.step(3)
    m = b.members[2]
    m.step(2)
    ref = Sim(fast=False, setup=setups(params)[2], verbose=False, rays=RAYS)
    ref.step(5)
    for name in ("rho", "mx", "my", "mz", "xi", "dfront"):
        same(getattr(m, name), getattr(ref, name), f"member 2 {name}")
    assert m.t == ref.t


def test_batch_plans_follow_live_parameters(torch_dev):
    """A member's live wind changed between batch steps reaches its stage plan (the plan key holds each member's
    wind, star count and table addresses): the member equals a Sim given the same change at the same step."""
    from nebulaviewer_amd.batch import SimBatch
    from nebulaviewer_amd.sim import Sim
    say("[batch] live wind change between steps")
    b = SimBatch(setups(), rays=RAYS, graph=True)
    b.step(2)
    b.members[3].set_param("wind", 0.2)
    b.step(3)
    ref = Sim(fast=False, setup=setups()[3], verbose=False, rays=RAYS)
    ref.step(2)
    ref.set_param("wind", 0.2)
    ref.step(3)
    for name in ("rho", "mx", "my", "mz", "xi"):
        same(getattr(b.members[3], name), getattr(ref, name), f"member 3 {name}")


def test_batch_dispatch_tuned_for_the_chosen_tile(torch_dev):
    """The dispatch form is measured with the tile the batch then runs (the tile is tuned first)."""
    from nebulaviewer_amd import autotune
    from nebulaviewer_amd.batch import SimBatch
    say("[batch] dispatch tuned after the tile")
    autotune._cache.clear()
    b = SimBatch(setups(), rays=RAYS)
    b.step(1)
    keys = [k for k in autotune.report() if "batch_dispatch" in k]
    assert len(keys) == 2
    for k in keys:
        stage = int(k.rstrip(")").split(",")[-1])
        assert f"'fp32', {b.tile[stage]}, {stage})" in k, (k, b.tile)
