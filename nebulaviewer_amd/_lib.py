"""ctypes binding of libnebula_hip.so (declared in include/nebula_hip.h).

The library is built in-tree by nebulaviewer_amd/build.py. Loading fails loudly when it is missing: there is
no CPU fallback behind the Sim class.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnebula_hip.so")

P = ctypes.c_void_p
I = ctypes.c_int
F = ctypes.c_float
LL = ctypes.c_longlong

# forms of the fused ionize pass (NV_ION_* in include/nebula_hip.h)
NV_ION_VEL = 1
NV_ION_FAST = 2

# entry point -> argument types (order and meaning as in include/nebula_hip.h)
SIGNATURES = {
    "nv_last_error": [],
    "nv_abi_version": [],
    "nv_hydro_stage": [P] * 13 + [I, I, I, P, I, F, P, F, F, F, F, F, F, F, F, F, F, F, I, I, P, P],
    "nv_hydro_stage_fast": [P] * 13 + [I, I, I, P, I, F, P, F, F, F, F, F, F, F, F, F, F, F, I, I, P, P],
    "nv_hydro_stage_h": [P] * 13 + [I, I, I, P, I, F, P, F, F, F, F, F, F, F, F, F, F, F, I, I, P, P],
    "nv_max_speed_h": [P] * 4 + [I, I, I, F, F, P, P],
    "nv_ionize_apply_h": [P, I, I, P, I, F, F, P, P, P, I, I, I, P, P, P, P, F, F, P, P],
    "nv_max_speed": [P] * 5 + [I, I, I, F, F, P, P],
    "nv_cast_rays": [P, I, I, I, F, F, F, F, F, I, F, F, F, F, F, P, I, I, P],
    "nv_cast_rays_wide": [P, I, I, I, F, F, F, F, F, I, F, F, F, F, F, P, I, I, P],
    "nv_ray_bricks_size": [I, I, I],
    "nv_ray_bricks": [P, I, I, I, P, P],
    "nv_ray_dirs": [I, I, P, P],
    "nv_ray_range_floats": [I, I],
    "nv_cast_rays_ex": [P, P, I, I, I, F, F, F, F, F, I, F, F, F, F, F, P, I, I, I, P],
    "nv_cast_rays_noalloc": [P, I, I, I, F, F, F, F, F, I, F, F, F, F, F, P, I, I, I, P],
    "nv_apply_front": [P, I, I, F, F, F, F, F, P, P, I, I, I, P],
    "nv_smooth_xi": [P, P, I, I, I, P],
    "nv_to_half": [P, P, LL, P],
    "nv_fill_f32": [P, F, LL, P],
    "nv_ionize_apply": [P, I, I, P, I, F, F, P, P, P, I, I, I, P, P, P, P, F, F, P, P],
    "nv_cast_fuv": [P, P, I, I, I, F, F, F, F, F, I, F, F, F, F, F, P, I, I, I, P],
    "nv_apply_fuv": [P, I, I, I, F, F, F, F, F, F, P, I, I, I, P],
    "nv_fuv_apply": [P, I, I, I, P, P, I, F, F, P, P, I, I, I, P],
    "nv_hydro_stage_batch_plan_size": [I],
    "nv_hydro_stage_ex": [P, I, I, I, I, P],
    "nv_hydro_stage_batch_plan": [P, I, I, I, I, I, P, LL],
    "nv_hydro_stage_batch": [P, P, P],
    "nv_hydro_stage_batch_queue": [P, P, P],
    "nv_hydro_stage_batch_queue_ex": [P, P, I, P],
    "nv_cfl_dt": [P, P, P, P, P, P, I, P],
    "nv_conservation_sums": [P] * 5 + [I, I, I, F, F, F, I, P, P],
    "nv_cast_rays_batch_plan_size": [I],
    "nv_cast_rays_batch_plan": [P, I, P, LL],
    "nv_cast_rays_batch": [P, P, P],
    "nv_ionize_apply_ex": [P, I, P],
    "nv_ionize_apply_batch_plan_size": [I],
    "nv_ionize_apply_batch_plan": [P, I, I, P, LL],
    "nv_ionize_apply_batch": [P, P, P],
    "nv_timeline_pack": [P] * 7 + [I] + [P] * 7 + [LL, P],
    "nv_timeline_unpack": [P] * 7 + [F] + [P] * 7 + [I, LL, P],
    "nv_shock_field": [P] * 4 + [I, I, I, F, P, P],
    "nv_frame_pack": [P] * 5 + [I] + [P] * 5 + [LL, P],
    "nv_pack_volume": [P] * 7 + [I, P, I, P, I, I, I, P],
    "nv_smooth_axis": [P, P, I, I, I, I, P],
    "nv_drive": [P] * 4 + [I, I, I] + [P] * 6 + [I, I, I, F, F, F, F, F, P],
    "nv_add_torus": [P] * 5 + [LL, P],
    "nv_mach_sums": [P] * 4 + [LL, P, P],
    "nv_every2": [P, I, P, I, I, I, P],
}

RESTYPES = {"nv_last_error": ctypes.c_char_p, "nv_hydro_stage_batch_plan_size": LL,
            "nv_cast_rays_batch_plan_size": LL, "nv_ionize_apply_batch_plan_size": LL, "nv_ray_bricks_size": LL,
            "nv_ray_range_floats": LL}


class NvStageDesc(ctypes.Structure):
    """one grid of nv_hydro_stage_batch_plan (include/nebula_hip.h)"""
    _fields_ = [("u", P * 4), ("xi", P), ("u0", P * 4), ("out", P * 4), ("stars", P), ("jets", P),
                ("smax_out", P), ("dt_dev", P), ("nx", I), ("ny", I), ("nz", I), ("nstars", I),
                ("wind", F), ("jet_cos", F), ("jet_r", F), ("pulse", F), ("dt", F), ("dx", F), ("a", F), ("b", F),
                ("floor_", F), ("cn2", F), ("ci2", F), ("vmax", F)]


class NvRayDesc(ctypes.Structure):
    """one job of nv_cast_rays_batch_plan"""
    _fields_ = [("rho", P), ("nx", I), ("ny", I), ("nz", I), ("sx", F), ("sy", F), ("sz", F), ("Q", F), ("dr", F),
                ("nsteps", I), ("dx", F), ("rcap", F), ("bx", F), ("by", F), ("bz", F), ("rfront", P), ("nu", I),
                ("nphi", I), ("wide", I), ("bricks", P), ("dirs", P), ("rf_range", P)]


class NvIonDesc(ctypes.Structure):
    """one grid of nv_ionize_apply_batch_plan"""
    _fields_ = [("rfronts", P), ("nu", I), ("nphi", I), ("star_xyz", P), ("nstars", I), ("dx", F), ("width", F),
                ("xi", P), ("dfront_h", P), ("dfront_f", P), ("nx", I), ("ny", I), ("nz", I), ("state", P * 4),
                ("cn2", F), ("ci2", F), ("smax_out", P), ("rf_ranges", P)]


class NvPackParams(ctypes.Structure):
    """render packing parameters (PackParams, render.py:77-88)"""
    _fields_ = [("c_ha", F * 3), ("c_o3", F * 3), ("c_s2", F * 3), ("c_dg", F * 3), ("c_sh", F * 3), ("gs", F),
                ("ambient", F), ("edge", F), ("inv_L", F * 3), ("dx", F)]


_lib = None


class NebulaHipError(RuntimeError):
    pass


def load(path=LIB_PATH):
    """Load the HIP library once and attach argument types. Raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise NebulaHipError(f"{path} is not built: run `python -m nebulaviewer_amd.build` (needs hipcc)")
    lib = ctypes.CDLL(path)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = RESTYPES.get(name, ctypes.c_int)
    if lib.nv_abi_version() != 5:
        raise NebulaHipError("libnebula_hip.so ABI version mismatch")
    _lib = lib
    return lib


def call(name, *args):
    """Call an entry point; raise NebulaHipError with the library's message on a nonzero status."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        msg = lib.nv_last_error().decode(errors="replace")
        raise NebulaHipError(f"{name} failed ({rc}): {msg}")
    return rc
