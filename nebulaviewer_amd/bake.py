"""The bake's on-disk frames: per frame `frames/f####_{rho,xi,df,g0,sh}.npy` and a closing `meta.json`, the format
the reference's bake writes (bake.py:94-120) and its Blender importer reads (blender/nebula_to_vdb.py:50-66).

    write_frame(sim, root, 0)            # one frame of the current state
    bake(sim, root, t_end=0.02, frames=40)

The stored forms (binary16 rho, dfront clipped to +-4, G0 and the shock proxy saturated at 6e4, uint8 xi) are
made on the GPU by nv_frame_pack in one pass; only the 9 bytes per cell of the frame cross PCIe. The scene
document, event replay and PNG previews of the reference's bake CLI are not part of this.
"""
import json
import os
import time

import numpy as np
import torch

from . import _lib

FRAME_FIELDS = ("g0", "rho", "xi", "df", "sh")


def frame_arrays(sim, shock=True):
    """the stored forms of the current state as host numpy arrays {name: array}"""
    sim.update_render_fields(force=True)
    sh = sim.shape
    dev = sim.device
    n = sh[0] * sh[1] * sh[2]
    h = lambda: torch.empty(sh, dtype=torch.float16, device=dev)
    o_rho, o_df, o_g0 = h(), h(), h()
    o_xi = torch.empty(sh, dtype=torch.uint8, device=dev)
    shock_t = sim.shock if shock else None
    o_sh = h() if shock else None
    half = 1 if sim.dfront.dtype == torch.float16 else 0
    rho = sim.rho
    _lib.call("nv_frame_pack", rho.data_ptr(), sim.xi.data_ptr(), sim.dfront.data_ptr(), sim.g0.data_ptr(),
              shock_t.data_ptr() if shock else None, half, o_rho.data_ptr(), o_xi.data_ptr(), o_df.data_ptr(),
              o_g0.data_ptr(), o_sh.data_ptr() if shock else None, n, sim._stream)
    out = dict(g0=o_g0, rho=o_rho, xi=o_xi, df=o_df)
    if shock:
        out["sh"] = o_sh
    return {k: v.cpu().numpy() for k, v in out.items()}


def write_frame(sim, root, frame, shock=True):
    """writes root/frames/f{frame:04d}_{name}.npy for each stored field; returns the paths"""
    d = os.path.join(root, "frames")
    os.makedirs(d, exist_ok=True)
    paths = []
    for name, arr in frame_arrays(sim, shock).items():
        p = os.path.join(d, f"f{frame:04d}_{name}.npy")
        np.save(p, arr)
        paths.append(p)
    return paths


def write_meta(sim, root, times, wall_s, **extra):
    """meta.json: the parameters plus grid, stars, jets and frame times (bake.py:108-111 keys)"""
    stars, jets = sim.stars_absolute()
    nu, nphi = (int(v) for v in sim.rfronts.shape[1:])
    meta = dict(sim.P, dx=sim.dx, shape=list(sim.shape), stars=stars, jets=jets, times=list(times), steps=sim.it,
                wall_s=wall_s, nu=nu, nphi=nphi, fuv_grid=list(sim.fuv_shape), generator="nebulaviewer_amd/bake.py")
    meta.update(extra)
    with open(os.path.join(root, "meta.json"), "w") as f:
        json.dump(meta, f, indent=1)
    return meta


def bake(sim, root, t_end, frames, shock=True, log=None):
    """steps the sim to t_end, writing `frames` evenly spaced frames (the first at the current time), each step
    capped so that a frame time is hit exactly (the schedule of bake.py:81-107)"""
    os.makedirs(root, exist_ok=True)
    t0 = time.perf_counter()
    dt_frame = t_end / max(frames - 1, 1)
    next_frame, frame, times = sim.t, 0, []
    while frame < frames:
        if sim.t >= next_frame - 1e-9:
            write_frame(sim, root, frame, shock)
            times.append(sim.t)
            if log:
                log(f"frame {frame}/{frames} t={sim.t:.4f} step {sim.it}")
            frame += 1
            next_frame += dt_frame
            continue
        sim.dt_cap = next_frame - sim.t + 1e-7
        sim.step(1)
    sim.dt_cap = None
    return write_meta(sim, root, times, time.perf_counter() - t0)
