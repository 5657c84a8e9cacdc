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
    "nv_apply_front": [P, I, I, F, F, F, F, F, P, P, I, I, I, P],
    "nv_smooth_xi": [P, P, I, I, I, P],
    "nv_to_half": [P, P, LL, P],
    "nv_fill_f32": [P, F, LL, P],
    "nv_ionize_apply": [P, I, I, P, I, F, F, P, P, P, I, I, I, P, P, P, P, F, F, P, P],
    "nv_cast_fuv": [P, P, I, I, I, F, F, F, F, F, I, F, F, F, F, F, P, I, I, I, P],
    "nv_apply_fuv": [P, I, I, I, F, F, F, F, F, F, P, I, I, I, P],
    "nv_fuv_apply": [P, I, I, I, P, P, I, F, F, P, P, I, I, I, P],
}

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
        fn.restype = ctypes.c_char_p if name == "nv_last_error" else ctypes.c_int
    if lib.nv_abi_version() != 1:
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
