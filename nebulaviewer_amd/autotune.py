"""Measured launch choices: the RK stage's workgroup tile, and whether the ray casts sample density bricks.

The stage kernel comes in three tile shapes (NV_TILE_* in include/nebula_hip.h): 64 z x 8 y and 32 z x 16 y
(8 wavefronts per workgroup, two workgroups per CU) and 64 z x 16 y (16 wavefronts, one workgroup per CU, half
the halo share), each marching its column over the library's depth (up to 64 x planes) or over the whole column
(NV_TILE_DEPTH). Shape and depth change the work decomposition only, never a bit of the result. Which one is fastest
depends on the z extent (idle lanes of the last z tile), on how many tiles the launch has to fill 256 CUs with,
and on how the halo share trades against occupancy, so it is measured rather than guessed: each candidate runs
the stage on the caller's own fields, timed with HIP events on the launch stream, and the fastest is kept. The
choice is cached per problem signature for the life of the process; the library's own rule (NV_TILE_AUTO, by idle
lanes) is what runs when tuning is switched off.

The ray casts can sample the density through its brick array (nv_ray_bricks, two 16-byte loads per trilinear
sample) or directly (four paired loads); both give the same fronts bit for bit. Bricks pay on large grids, where
the casts are bound by the cache lines their gathers touch (512^3: 3.2 ms of casts become 1.8 ms plus a 0.5 ms
build), and lose on small ones, whose density stays in the caches anyway (the 64-member batch of 48^3-160^3 grids:
2.7 ms direct, 3.0 ms with bricks), so that choice is measured too, on the first ionization.

    t1, t2 = tune_sim(sim)               # Sim(tile="tune") does this on its first RK stage
    tiles = tune_batch(batch, pulses)    # SimBatch(tile="tune") does this when it builds its stage plans
    use = tune_rays(sim, nstars, casts)  # Sim(ray_bricks="auto") on its first ionization
    report()                             # {signature: {"choice": name, "ms": {name: ms, ...}}}
"""
import ctypes

import torch

from . import _lib

TILES = {1: "64x8", 2: "32x16", 3: "64x16"}
CANDIDATES = (1, 2, 3)
PRECISION_CODES = {"fp32": 0, "fp32-fast": 1, "fp16-vel": 2}
WHOLE_COLUMN = 1 << 15   # NV_TILE_DEPTH(n) with n >= every grid's nx: each workgroup marches its whole column


def depth_code(shape, depth):
    """the NV_TILE_* code of a tile shape (1-3) with a march depth (0: the library's choice)"""
    return shape | (depth << 8)


def tile_name(code):
    """'64x8', or '64x8/whole' / '64x8/d<n>' when the code carries a march depth"""
    shape, depth = code & 0xFF, code >> 8
    name = TILES.get(shape, str(shape))
    return name if depth == 0 else name + ("/whole" if depth >= WHOLE_COLUMN else f"/d{depth}")


# each shape with the library's march depth and with whole-column marches (one workgroup per tile column: the
# columns of one XCD's run advance together, so the halo lines they share are read while still in its L2)
SIM_CANDIDATES = tuple(depth_code(t, d) for t in CANDIDATES for d in (0, WHOLE_COLUMN))

_cache = {}


NAMES = {"sim": tile_name, "batch": tile_name, "rays": {False: "direct", True: "bricks"}.get,
         "batch_rays": {False: "direct", True: "bricks"}.get,
         "batch_dispatch": lambda c: "hardware" if not c[0] else ("queue" if c[1] == 0 else f"queue-{c[1]}wg/CU")}


def report():
    """every tuning result of this process: {signature: {"choice": name, "ms": {candidate name: ms per launch}}}"""
    out = {}
    for k, v in _cache.items():
        name = NAMES[k[0]]
        out[str(k)] = {"choice": name(v["choice"]), "ms": {name(c): round(ms, 4) for c, ms in v["ms"].items()}}
    return out


def _time(launch, stream_dev, reps, warm=2):
    """mean ms per call of launch() over reps calls, after warm untimed calls (HIP events on the launch stream)"""
    strm = torch.cuda.current_stream(stream_dev)
    for _ in range(warm):
        launch()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record(strm)
    for _ in range(reps):
        launch()
    e.record(strm)
    e.synchronize()
    return s.elapsed_time(e) / reps


def _pick(key, measure, candidates, rounds=2):
    """the candidate with the least mean time; candidates are measured in interleaved rounds, so that a drift of
    the device's speed over the first launches of a process (clocks, first touches) does not favour the later ones"""
    if key in _cache:
        return _cache[key]["choice"]
    ms = {t: 0.0 for t in candidates}
    for _ in range(rounds):
        for t in candidates:
            ms[t] += measure(t) / rounds
    best = min(ms, key=ms.get)
    _cache[key] = {"choice": best, "ms": ms}
    return best


def tune_sim(sim, candidates=SIM_CANDIDATES, reps=5):
    """The fastest tile of each stage form for this Sim's grid, precision, solver and limiter: (stage-1 tile,
    stage-2 tile). Stage 1 is timed as the step runs it (U -> the free scratch block sim.stage); stage 2 with the
    step's traffic (reads U as U1 and the scratch block as U0, writes the scratch block in place, fused CFL
    maximum into a scratch float), so the simulation state is untouched."""
    solver, lim = sim._solver_lim()
    keys = [("sim", tuple(sim.shape), sim.precision, solver, lim, st) for st in (1, 2)]
    if all(k in _cache for k in keys):
        return tuple(_cache[k]["choice"] for k in keys)
    code = PRECISION_CODES[sim.precision]
    scratch = torch.zeros(1, dtype=torch.float32, device=sim.device)
    U = [t.data_ptr() for t in sim.U]
    U1 = [sim.stage[c].data_ptr() for c in range(4)]
    descs = {1: sim._stage_desc(U, U, U1, 0.0, 1.0, 0.0, len(sim.stars), sim.P["wind"], 1.0, None),
             2: sim._stage_desc(U, U1, U1, 0.5, 0.5, 0.0, len(sim.stars), sim.P["wind"], 1.0, scratch.data_ptr())}
    out = []
    for st, key in zip((1, 2), keys):
        d = descs[st]

        def measure(tile, d=d):
            strm = sim._stream
            return _time(lambda: _lib.call("nv_hydro_stage_ex", ctypes.byref(d), code, solver, lim, tile, strm),
                         sim.device, reps)

        out.append(_pick(key, measure, candidates))
    return tuple(out)


def tune_batch(batch, pulses, candidates=SIM_CANDIDATES, reps=3):
    """The fastest tile of each stage form for a SimBatch's member grids: {1: tile, 2: tile}. Stage 1 is timed on
    its own plan (it writes only the members' scratch blocks); stage 2 on a plan with the step's traffic (U as U1,
    the scratch blocks as U0) whose outputs go to a scratch buffer, so the state is untouched."""
    shapes = tuple(tuple(m.shape) for m in batch.members)
    keys = {st: ("batch", shapes, batch.precision, batch.solver, batch.lim, st) for st in (1, 2)}
    if all(k in _cache for k in keys.values()):
        return {st: _cache[k]["choice"] for st, k in keys.items()}
    bufs, scratch = _scratch_outputs(batch)
    smax = torch.zeros(batch.n, dtype=torch.float32, device=batch.device)
    out = {}
    for st, key in keys.items():
        def measure(tile, st=st):
            batch.tile = {1: tile, 2: tile}
            plan = batch._stage_plan(st, pulses, scratch=scratch if st == 2 else None,
                                     smax=smax.data_ptr() if st == 2 else None)
            return _time(lambda: _lib.call("nv_hydro_stage_batch", *plan.args, batch._stream), batch.device, reps)

        out[st] = _pick(key, measure, candidates)
    torch.cuda.synchronize(batch.device)
    del bufs
    batch._stage_key = None
    return out


def _scratch_outputs(batch):
    """per member, four output pointers into a scratch buffer of 16 B per cell (float32 fields at 4 B, binary16
    velocities at 2 B per cell); the buffers are returned so the caller keeps them alive"""
    bufs = [torch.empty(4 * m.shape[0] * m.shape[1] * m.shape[2], dtype=torch.float32, device=batch.device)
            for m in batch.members]
    scratch = []
    for b, m in zip(bufs, batch.members):
        n = b.numel() // 4
        base = b.data_ptr()
        scratch.append([base, base + 4 * n, base + 8 * n, base + 12 * n] if not m.vel_state else
                       [base, base + 4 * n, base + 6 * n, base + 8 * n])
    return bufs, scratch


def tune_rays(sim, nstars, casts, reps=2):
    """True when the brick-sampled casts (brick build included) beat the direct ones for this grid, ray grid and
    number of active stars. casts(use_bricks) launches one ionization's casts; both forms write the same fronts."""
    key = ("rays", tuple(sim.shape), tuple(sim.rfronts.shape[1:]), nstars)
    return _pick(key, lambda use: _time(lambda: casts(use), sim.device, reps), (False, True))


def tune_batch_rays(batch, plans, reps=2):
    """True when the batch's brick-sampled ray plan beats the direct one; plans = {False: plan, True: plan}"""
    key = ("batch_rays", tuple(tuple(m.shape) for m in batch.members), tuple(len(batch._on(m)) for m in batch.members))

    def measure(use):
        plan = plans[use]
        return _time(lambda: _lib.call("nv_cast_rays_batch", *plan.args, batch._stream), batch.device, reps)
    return _pick(key, measure, (False, True))


# how a stage's tiles reach the CUs: (persistent queue?, its resident workgroups per CU, 0 = as many as fit)
DISPATCH_CANDIDATES = ((False, 0), (True, 0), (True, 1))


def tune_batch_dispatch(batch, pulses, reps=3, candidates=DISPATCH_CANDIDATES):
    """{1: (use_queue, wgs_per_cu), 2: ...}: for each stage form, whether the persistent tile-queue launch
    (nv_hydro_stage_batch_queue_ex) beats one workgroup per tile, and at which residency: the waves per CU of the
    persistent grid are its workgroups per CU times the tile's wavefronts (16 at full residency for every tile
    shape, 8 with one 512-thread workgroup per CU). Timed on plans whose outputs go to a scratch buffer (16 B per
    cell and member, freed afterwards), so the state is untouched; the queue form exists for HLL + minmod only."""
    if (batch.solver, batch.lim) != (1, 0):
        return {1: (False, 0), 2: (False, 0)}
    shapes = tuple(tuple(m.shape) for m in batch.members)
    out = {}
    bufs, scratch = _scratch_outputs(batch)
    smax = torch.zeros(batch.n, dtype=torch.float32, device=batch.device)
    for stage in (1, 2):
        key = ("batch_dispatch", shapes, batch.precision, batch.tile[stage], stage)
        plan = batch._stage_plan(stage, pulses, scratch=scratch, smax=smax.data_ptr())

        def measure(c, plan=plan):
            if c[0]:
                return _time(lambda: _lib.call("nv_hydro_stage_batch_queue_ex", *plan.args, c[1], batch._stream),
                             batch.device, reps)
            return _time(lambda: _lib.call("nv_hydro_stage_batch", *plan.args, batch._stream), batch.device, reps)
        out[stage] = _pick(key, measure, candidates)
    torch.cuda.synchronize(batch.device)
    del bufs
    return out
