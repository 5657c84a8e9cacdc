"""The renderer's volume packing on the step's fields: RGBA16F emission + dust per cell (nv_pack_volume, the
reference's render.pack_volume_lm, render.py:176-204) and the optional [1 4 1]/6 prefilter along each axis
(nv_smooth_axis, render.py:281-298). The ray marcher, cameras, palettes and tone mapping are the renderer's and
are not part of this.

    pp = pack_params(sim, c_ha=(1, .2, .1), c_o3=(0, .6, .5), c_s2=(.8, 0, 0), c_dg=(.1, .1, .1), c_sh=(0, 0, 0),
                     gs=10.0, ambient=0.02, edge=0.05)
    vol = pack_volume(sim, pp, smooth=True)      # (nx, ny, nz, 4) float16
"""
import torch

from . import _lib


def pack_params(sim, c_ha, c_o3, c_s2, c_dg, c_sh, gs, ambient, edge):
    """NvPackParams (PackParams, render.py:77-88) with the sim's geometry: inv_L = 1 / box size, dx"""
    pp = _lib.NvPackParams()
    for name, v in (("c_ha", c_ha), ("c_o3", c_o3), ("c_s2", c_s2), ("c_dg", c_dg), ("c_sh", c_sh)):
        getattr(pp, name)[:] = [float(c) for c in v]
    pp.gs, pp.ambient, pp.edge = float(gs), float(ambient), float(edge)
    pp.inv_L[:] = [1.0 / v for v in sim.L]
    pp.dx = float(sim.dx)
    return pp


def pack_volume(sim, pp, smooth=False):
    """the packed RGBA16F volume (nx, ny, nz, 4) of the sim's current fields; the shock term is computed in the
    pack only when c_sh is non-zero (render.py:734-741)"""
    import ctypes
    sim.update_render_fields()
    sh = tuple(sim.shape)
    dev = sim.device
    a = torch.empty(sh + (4,), dtype=torch.float16, device=dev)
    half = 1 if sim.dfront.dtype == torch.float16 else 0
    shock_on = 1 if any(c != 0.0 for c in pp.c_sh) else 0
    rho, mx, my, mz = sim.rho, sim.mx, sim.my, sim.mz
    strm = sim._stream
    _lib.call("nv_pack_volume", rho.data_ptr(), mx.data_ptr(), my.data_ptr(), mz.data_ptr(), sim.xi.data_ptr(),
              sim.dfront.data_ptr(), sim.g0.data_ptr(), half, ctypes.byref(pp), shock_on, a.data_ptr(), *sh, strm)
    if not smooth:
        return a
    b = torch.empty_like(a)
    for axis, (src, dst) in enumerate(((a, b), (b, a), (a, b))):
        _lib.call("nv_smooth_axis", src.data_ptr(), dst.data_ptr(), axis, *sh, strm)
    return b
