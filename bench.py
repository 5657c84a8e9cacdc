"""Throughput of the simulation step on one MI355X: cell-updates/s and the HBM roofline of the hydro kernel.

python bench.py [--steps K] [--warmup W] [--config 2] [--no-cpu]

Workload (default): the 512^3 ionized cliff wall (n=100 wall with a tanh edge, lognormal sigma 0.4, 30 dense
globules, one star 0.4 above the face), FP32, HLL + minmod MUSCL, RK2, CFL 0.3, ionization every 2 steps with
1024x2048 rays at half-voxel steps. Inputs are generated on the host before timing and resident in HBM.

One step = Sim.step(1): the ray casts and the fused ionize pass on ionization steps, the CFL readback, and the
two fused RK stages. value = cells * K / wall time of K steps (ionization amortised over its cadence).

roofline: the dominant kernel (the fused RK stage). Algorithmic bytes per cell per RK2 step: 88 (stage 1 reads
U + xi and writes U1: 36 B; stage 2 reads U1 + xi + U^n and writes U^n+1: 52 B), i.e. 44 B per cell per launch
on average; achieved = 44 B x cells / (mean launch time from HIP events on the launch stream).
ionization_roofline: the ray casts and the fused ionize pass, each against its own algorithmic bytes per cell and
call (4 B per active star for the casts' density read; 22 B for the ionize pass: xi and dfront written, the state
read for the fused CFL max).
cpu_baseline: the C++ oracle (oracle/, a restatement of the reference kernels; the reference's Warp CPU device
cannot run here) on the same 512^3 grid: one ionization and one hydro step on all host cores.
variants: the same workload in the fast-arithmetic and binary16-velocity forms, and the uneven batch (64
independent members of 48^3..160^3, scenarios.batch_setups) stepped by SimBatch with HIP-graph replay, in the
headline's form ("batch") and in the binary16-velocity form ("batch_fp16-vel"), each with its throughput relative
to the single 512^3 grid of the same form (full step and hydro only).
"""
import argparse
import contextlib
import gc
import json
import math
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# algorithmic bytes per cell per RK2 step: 5 S_U + 2 S_xi (S_xi = 4): S_U = 16 for the float32 state,
# 10 for float32 density with binary16 velocity
BYTES_PER_CELL_STEP = {"fp32": 88.0, "fp32-fast": 88.0, "fp16-vel": 58.0}
HEADLINE = "fp32-fast"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


@contextlib.contextmanager
def no_gc():
    """Python's cyclic garbage collector collected before the timed steps and paused during them, as timeit does:
    a full collection in a process holding torch takes tens of milliseconds on the host, and one landing between
    two launches leaves the device idle. Before this, three of five runs had one ray-cast window of the first form's
    50 steps stretched from 1.2 ms to 7-23 ms, with every other launch at its usual time. Three runs with the
    collector paused had none."""
    gc.collect()
    was = gc.isenabled()
    gc.disable()
    try:
        yield
    finally:
        if was:
            gc.enable()


def measured_traffic(precision):
    """HBM bytes per launch of the stage kernel of this precision form on the 512^3 wall, from the committed
    rocprofv3 --pmc summary (FETCH_SIZE and WRITE_SIZE runs scaled by a plain copy's counts); null if absent."""
    path = os.path.join(ROOT, "profiles", "hydro_stage_wall512_traffic.md")
    if not os.path.exists(path):
        return None
    m = re.search(r"hbm_bytes_per_launch\[" + re.escape(precision) + r"\]:\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


def host_cores():
    """the host's CPU inventory: os.cpu_count(), the CPUs this process may run on, the OpenMP thread count the
    environment grants, the CPU model; `use` is the thread count the CPU baseline runs with (all granted CPUs)"""
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return dict(os_cpu_count=os.cpu_count(), affinity=aff, omp_num_threads=omp or None, model=model,
                use=min(aff, omp) if omp > 0 else aff)


def cpu_baseline(setup, ion_every):
    import numpy as np
    from nebulaviewer_amd import scenarios as S
    from oracle.oracle import OracleSim
    host = host_cores()
    threads = host["use"]
    C = S.constants(setup)
    t0 = time.perf_counter()
    ref = OracleSim(setup, C, S.ray_grid(setup.shape[2]), S.fuv_grid(setup.shape[2]), np.float32, threads=threads)
    t_ion = time.perf_counter() - t0          # construction = the first ionize() (casts + per-cell passes)
    ref.it = 1                                # next step is not an ionization step: time dt + 2 stages alone
    t0 = time.perf_counter()
    ref.step(1)
    t_hydro = time.perf_counter() - t0
    n = math.prod(setup.shape)
    per_step = t_hydro + t_ion / ion_every
    # one host thread (the reference's CPU launch semantics), bounded sample: both RK stages on an x-slab of the
    # same state (the stage cost per cell does not depend on the slab's thickness beyond its 2-plane rim)
    from oracle.oracle import Kernels
    k1 = Kernels(np.float32, threads=1)
    nx1 = min(setup.shape[0], max(8, (1 << 22) // (setup.shape[1] * setup.shape[2])))
    sl = (nx1,) + tuple(setup.shape[1:])
    U = [np.ascontiguousarray(f[:nx1]) for f in (ref.rho, ref.mx, ref.my, ref.mz)]
    xi = np.ascontiguousarray(ref.xi[:nx1])
    U1 = [np.empty_like(u) for u in U]
    P, C = ref.P, ref.C
    hll = P.get("solver", "rusanov") == "hll"
    lim = {"minmod": 0, "vanleer": 1, "mc": 2}[P.get("limiter", "minmod")] if hll else 0
    common = dict(shape=sl, stars=ref.star_arr, nstars=len(ref.stars), wind=P["wind"], jets=ref.jet_arr,
                  jet_cos=C["jet_cos"], jet_r=P["jet_r"], pulse=1.0, dt=ref.dt or 1e-4, dx=ref.dx, floor=P["floor"],
                  cn2=C["cn2"], ci2=C["ci2"], vmax=6.0, solver=int(hll), lim=lim)
    t0 = time.perf_counter()
    k1.hydro_stage(U, xi, U, U1, a=0.0, b=1.0, **common)
    k1.hydro_stage(U1, xi, U, U, a=0.5, b=0.5, **common)
    t1 = time.perf_counter() - t0
    single = math.prod(sl) / t1
    return dict(value=n / per_step, unit="cell-updates/s", cores=threads, kind="port", host=host,
                sample=f"C++ oracle (float, -O3, OpenMP {threads} threads) on the same {setup.shape[0]}^3 grid: one "
                       f"ionization ({t_ion:.2f} s) and one CFL + two-stage hydro step ({t_hydro:.2f} s), "
                       f"per step = hydro + ionization/{ion_every}",
                single_thread={"value": single, "unit": "cell-updates/s", "cores": 1, "scope": "hydro (2 RK stages)",
                               "sample": f"both RK stages of the C++ oracle on one host thread over a {sl[0]}x{sl[1]}x"
                                         f"{sl[2]} x-slab of the same state ({t1:.2f} s)"})


def ionization_roofline(ks, cells, nstars, precision):
    """HBM roofline of the ionization kernels, each against its own algorithmic bytes per cell and call: the
    brick build reads the density and writes its 16-byte bricks (20 B), the ray casts read the density once per
    active star (4 B per star; the gathers' re-reads are not counted),
    the fused ionize pass writes xi (4 B) and dfront (2 B) and reads the state for the fused CFL max (16 B, or
    10 B with binary16 velocity); its unsmoothed xi planes stay in LDS."""
    per_cell = {"ray_bricks_kernel": 20.0, "cast_rays_kernel": 4.0 * nstars,
                "ionize_apply_kernel": 6.0 + (10.0 if precision == "fp16-vel" else 16.0)}
    out = {"bound": "hbm", "unit": "GB/s", "peak": HBM_PEAK_GBS, "stars": nstars, "kernels": {}}
    for key, name in (("ray_bricks", "ray_bricks_kernel"), ("cast_rays", "cast_rays_kernel"),
                      ("ionize_apply", "ionize_apply_kernel")):
        if key not in ks:
            continue
        ms = ks[key]["avg_ms"]
        ach = per_cell[name] * cells / (ms / 1e3) / 1e9
        out["kernels"][name] = {"ms": round(ms, 4), "bytes_per_cell": per_cell[name], "achieved": ach,
                                "frac": ach / HBM_PEAK_GBS}
    return out


def mass_budget(sim):
    """One more step with the FP64 conservation sums on the GPU (outside the timed region): the change of total
    mass M against the boundary mass flux of the two RK stages, M^{n+1} - M^n = -(dt/dx) (B(U^n) + B(U1)) / 2,
    which holds to the rounding of the cells' float updates while the density floor does not act."""
    import torch
    import ctypes
    from nebulaviewer_amd import _lib
    f = ctypes.c_float

    def sums(U):
        o = torch.zeros(2, dtype=torch.float64, device=sim.device)
        _lib.call("nv_conservation_sums", *[t.data_ptr() for t in U], sim.xi.data_ptr(), *sim.shape,
                  f(sim.P["floor"]), f(sim.cn2), f(sim.ci2), 1 if sim.P.get("solver", "rusanov") == "hll" else 0,
                  o.data_ptr(), sim._stream)
        return o.tolist()

    m0, b0 = sums(sim.U)
    rho0 = sim.rho.double()
    bound = 12.0 * float(torch.sqrt(torch.sum((rho0.abs() * 2.0 ** -23) ** 2)))   # ~ the cells' float roundings
    floor_on = float(sim.rho.min()) <= 1.0001 * sim.P["floor"]
    del rho0
    sim.step(1)
    m1, b1 = sums([sim.stage[c] for c in range(4)])
    m2 = sums(sim.U)[0]
    flux = -0.5 * float(f(sim.dt).value) / sim.dx * (b0 + b1)
    return {"dM": m2 - m0, "boundary_flux_term": flux, "residual": (m2 - m0) - flux, "rounding_bound": bound,
            "M": m0, "floor_active": floor_on}


def copy_calibration(nbytes):
    """A plain copy kernel of known size (reads and writes nbytes each launch), to scale the HBM byte counters."""
    import torch
    n = nbytes // 4
    src = torch.rand(n, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / 3
    print(json.dumps({"copy_bytes": n * 4, "ms": dt * 1e3, "GBps_read_plus_write": 2 * n * 4 / dt / 1e9}), flush=True)


def sustained_copy():
    """What this device sustains for a plain stream, measured in the same run: a 1 GiB device-to-device copy
    (torch's copy kernel), timed with HIP events over 10 launches after 2 untimed ones; GB/s of bytes read plus
    bytes written. The published 8 TB/s is the roofline's peak; this is the ceiling a streaming kernel meets here."""
    import torch
    n = (1 << 30) // 4
    src = torch.rand(n, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(2):
        dst.copy_(src)
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(10):
        dst.copy_(src)
    e.record()
    e.synchronize()
    ms = s.elapsed_time(e) / 10
    del src, dst
    torch.cuda.empty_cache()
    return {"kernel": "1 GiB device copy (read + write)", "ms": round(ms, 4), "GBps": 2 * n * 4 / (ms * 1e-3) / 1e9}


def make_measure_batch(a, S, KernelTimer):
    import torch
    from nebulaviewer_amd import autotune

    made = {}

    def measure_batch(count, precision=None):
        from nebulaviewer_amd.batch import SimBatch
        precision = precision or a.precision
        if count not in made:
            t0 = time.perf_counter()
            made[count] = S.batch_setups(count)
            log(f"[bench] batch of {count}: initial conditions in {time.perf_counter() - t0:.1f} s")
        setups = made[count]
        b = SimBatch(setups, precision=precision, graph=True)
        cells = b.cells
        b.step(a.warmup)                       # captures the two kinds of step
        torch.cuda.synchronize()
        with no_gc():
            t0 = time.perf_counter()
            b.step(a.steps)
            torch.cuda.synchronize()
            wall = time.perf_counter() - t0
        timer = KernelTimer(reserve=32 * b.ion_every)
        b.timer = timer                        # per-kernel events (eager launches, no graph)
        b.step(2 * b.ion_every)
        torch.cuda.synchronize()
        b.timer = None
        ks = timer.summary()
        hyd_s = ks["hydro_stage"]["avg_ms"] / 1e3
        r = dict(members=count, cells=cells, stars=sum(len(m.stars) for m in b.members),
                 sizes=sorted({m.shape[0] for m in b.members}), value=cells * a.steps / wall,
                 ms_per_step=wall / a.steps * 1e3, hydro_launch_ms=hyd_s * 1e3,
                 hydro_only_cell_updates_per_s=cells / (2 * hyd_s),
                 hydro_achieved_GBps=BYTES_PER_CELL_STEP[precision] / 2 * cells / hyd_s / 1e9, precision=precision,
                 kernels_ms={k: round(v["avg_ms"], 4) for k, v in ks.items()}, graph_replay=True,
                 tile={f"stage{k}": autotune.tile_name(t) for k, t in b.tile.items()},
                 ray_bricks=getattr(b, "ray_bricks_used", None),
                 dispatch={f"stage{k}": (f"queue, {b.queue_wgs[k] or 'all'} workgroups/CU" if v else "hardware")
                           for k, v in b.dispatch.items()})
        log(f"[bench] batch ({precision}): {r['value'] / 1e9:.2f} G cell-updates/s ({cells / 1e6:.1f} M cells, {r['stars']} stars), "
            f"{r['ms_per_step']:.3f} ms/step, hydro {hyd_s * 1e3:.3f} ms/launch; kernels {r['kernels_ms']}")
        del b
        torch.cuda.empty_cache()
        return r

    return measure_batch


def batch_only(a):
    """--config 4: the uneven batch as the measured workload (one JSON line of its own)"""
    from nebulaviewer_amd import autotune
    from nebulaviewer_amd import scenarios as S
    from nebulaviewer_amd.timing import KernelTimer
    r = make_measure_batch(a, S, KernelTimer)(a.batch or 64)
    out = {
        "metric": f"cell-updates/s, batch of {r['members']} uneven grids ({a.precision} step)",
        "value": r["value"], "unit": "cell-updates/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "precision": a.precision, "data": "synthetic (seeded initial conditions generated on the host)",
        "config": {"workload": f"batch of {r['members']} members (torus / sheet, N in {r['sizes']}), HLL+minmod "
                               "RK2, CFL 0.3, ionization every 2 steps with 512x1024 rays, seeds 100+i",
                   "cells": r["cells"], "stars": r["stars"]},
        "roofline": {"bound": "hbm", "kernel": "hydro_stage_batch_kernel", "achieved": r["hydro_achieved_GBps"],
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": r["hydro_achieved_GBps"] / HBM_PEAK_GBS,
                     "traffic": None, "algorithmic_bytes_per_launch": BYTES_PER_CELL_STEP[a.precision] / 2 * r["cells"]},
        "hydro_only_cell_updates_per_s": r["hydro_only_cell_updates_per_s"], "kernels_ms": r["kernels_ms"],
        "tile": r["tile"], "ray_bricks": r["ray_bricks"], "dispatch": r["dispatch"], "autotune": autotune.report(),
        "variants": {}, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--config", type=int, default=2, help="0-3: one grid of scenarios.CONFIGS; 4: the uneven batch")
    ap.add_argument("--n", type=int, default=None, help="override the grid size (testing only)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--precision", default=HEADLINE, choices=["fp32", "fp32-fast", "fp16-vel"],
                    help="the form whose throughput is `value`: fp32-fast (FP32 arithmetic with fused multiply-adds "
                         "and hardware reciprocals, within the drop-in tolerance of the FP64 oracle), fp32 (the "
                         "reference's operation order, bit-exact against the CPU oracle) or fp16-vel")
    ap.add_argument("--no-variants", action="store_true", help="do not also time the other two forms")
    ap.add_argument("--batch", type=int, default=64,
                    help="members of the uneven-batch variant (0: skip it; timed with the variants)")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="counter calibration: run a plain device-to-device copy of this many bytes 3 times, exit")
    a = ap.parse_args()
    if a.copy_bytes:
        copy_calibration(a.copy_bytes)
        return
    if a.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {a.gpus} is not supported (the step has no "
              "multi-GPU path)", file=sys.stderr)
        sys.exit(2)

    import torch
    from nebulaviewer_amd import scenarios as S
    from nebulaviewer_amd.sim import Sim
    from nebulaviewer_amd.timing import KernelTimer
    from nebulaviewer_amd import autotune

    if a.config == 4:
        batch_only(a)
        return
    scenario, n, seed = S.CONFIGS[a.config]
    n = a.n or n
    t0 = time.perf_counter()
    setup = S.make_setup(scenario, n, seed)
    log(f"[bench] {scenario} {n}^3 initial conditions in {time.perf_counter() - t0:.1f} s")
    def measure(precision):
        sim = Sim(fast=False, setup=setup, verbose=False, precision=precision)
        log(f"[bench] {precision}: warmup {a.warmup} steps, timing {a.steps} steps")
        sim.step(a.warmup)
        torch.cuda.synchronize()
        timer = KernelTimer(reserve=16 * a.steps)   # every event of the timed steps created before them
        sim.timer = timer
        torch.cuda.synchronize()
        with no_gc():
            t0 = time.perf_counter()
            sim.step(a.steps)
            torch.cuda.synchronize()
            wall = time.perf_counter() - t0
        sim.timer = None
        ks = timer.summary()
        cells = math.prod(sim.shape)
        hyd_s = ks["hydro_stage"]["avg_ms"] / 1e3
        achieved = BYTES_PER_CELL_STEP[precision] / 2 * cells / hyd_s / 1e9
        r = dict(value=cells * a.steps / wall, ms_per_step=wall / a.steps * 1e3, hydro_launch_ms=hyd_s * 1e3,
                 hydro_only_cell_updates_per_s=cells / (2 * hyd_s), hydro_achieved_GBps=achieved,
                 kernels_ms={k: round(v["avg_ms"], 4) for k, v in ks.items()},
                 kernels_min_max_ms={k: (round(v["min_ms"], 4), round(v["max_ms"], 4)) for k, v in ks.items()},
                 ion_every=sim.ion_every,
                 rays=tuple(sim.rfronts.shape[1:]), nstars=len(sim.stars),
                 tile={f"stage{k + 1}": autotune.tile_name(t) for k, t in enumerate(sim.tile_shape)},
                 ray_bricks=getattr(sim, "ray_bricks_used", None),
                 ionization_roofline=ionization_roofline(ks, cells, sum(1 for st in sim.stars if sim.t >= st[4]),
                                                         precision))
        log(f"[bench] {precision}: {r['value'] / 1e9:.2f} G cell-updates/s, {r['ms_per_step']:.3f} ms/step, "
            f"hydro {hyd_s * 1e3:.3f} ms/launch = {achieved:.0f} GB/s; kernels {r['kernels_ms']}, "
            f"(least, greatest) {r['kernels_min_max_ms']}")
        if precision == "fp32":
            r["mass_budget"] = mass_budget(sim)
            log(f"[bench] {precision}: mass budget of one more step {r['mass_budget']}")
        del sim
        torch.cuda.empty_cache()
        return r

    measure_batch = make_measure_batch(a, S, KernelTimer)

    main_r = measure(a.precision)
    variants = {}
    if not a.no_variants:
        for p in ("fp32", "fp32-fast", "fp16-vel"):
            if p != a.precision:
                variants[p] = measure(p)
        if a.batch > 0:
            # the batch in the headline's precision and in the reduced-storage form, each against the single grid
            # of the same form
            for p, key in ((a.precision, "batch"), ("fp16-vel", "batch_fp16-vel")):
                if key != "batch" and p == a.precision:
                    continue
                br = measure_batch(a.batch, p)
                ref = main_r if p == a.precision else variants[p]
                br["vs_single_grid"] = {"step": br["value"] / ref["value"],
                                        "hydro_only": br["hydro_only_cell_updates_per_s"] /
                                        ref["hydro_only_cell_updates_per_s"]}
                variants[key] = br
    cells = math.prod(setup.shape)
    value = main_r["value"]
    achieved = main_r["hydro_achieved_GBps"]
    sustained = sustained_copy()
    log(f"[bench] sustained stream: {sustained['GBps']:.0f} GB/s ({sustained['kernel']})")
    # the committed counter profile covers the FP32 stage on the 512^3 wall only
    traffic = measured_traffic(a.precision) if (a.config == 2 and n == 512) else None

    cpu = None
    if not a.no_cpu:
        log("[bench] CPU baseline (C++ oracle)")
        cpu = cpu_baseline(setup, main_r["ion_every"])
        log(f"[bench] CPU baseline {cpu['value'] / 1e6:.2f} M cell-updates/s")

    out = {
        "metric": (f"cell-updates/s at {n}^3 (FP32 step) and % of HBM peak" if a.precision != "fp16-vel"
                   else f"cell-updates/s at {n}^3 (FP16-storage step) and % of HBM peak"),
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": main_r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "precision": a.precision,
        "data": "synthetic (seeded initial conditions generated on the host)",
        "config": {"workload": f"{scenario} {n}^3, HLL+minmod RK2, CFL 0.3, ionization every "
                               f"{main_r['ion_every']} steps with {main_r['rays'][0]}x{main_r['rays'][1]} rays "
                               f"(dr = dx/2), seed {seed}", "cells": cells, "stars": main_r["nstars"]},
        "roofline": {"bound": "hbm", "kernel": "hydro_stage_kernel", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic,
                     "algorithmic_bytes_per_launch": BYTES_PER_CELL_STEP[a.precision] / 2 * cells,
                     "sustained": sustained, "frac_of_sustained": achieved / sustained["GBps"]},
        "hydro_only_cell_updates_per_s": main_r["hydro_only_cell_updates_per_s"],
        "kernels_ms": main_r["kernels_ms"],
        "ionization_roofline": main_r["ionization_roofline"],
        "mass_budget": main_r.get("mass_budget") or variants.get("fp32", {}).get("mass_budget"),
        "autotune": autotune.report(),
        "variants": variants,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
