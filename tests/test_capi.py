"""The C ABI library: it loads, exports every entry point include/nebula_hip.h declares with the argument count
the ctypes binding uses, and rejects bad arguments with a message before touching the GPU."""
import ctypes
import os
import re

import pytest

from nebulaviewer_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "nebula_hip.h")


def declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    out = {}
    for m in re.finditer(r"\b(?:int|const char\*)\s+(nv_\w+)\s*\(([^)]*)\)\s*;", src):
        args = m.group(2).strip()
        out[m.group(1)] = 0 if args in ("", "void") else args.count(",") + 1
    return out


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libnebula_hip.so is not built (run __graft_entry__.build())")
    return _lib.load()


def test_every_declared_symbol_is_exported_and_bound(lib):
    decl = declared()
    assert len(decl) >= 13
    for name, nargs in decl.items():
        assert hasattr(lib, name), f"{name} not exported"
        assert name in _lib.SIGNATURES, f"{name} has no ctypes binding"
        assert len(_lib.SIGNATURES[name]) == nargs, f"{name}: header has {nargs} args"
    assert set(_lib.SIGNATURES) == set(decl)


def test_abi_version(lib):
    assert lib.nv_abi_version() == 1


def test_rejects_null_fields(lib):
    f = ctypes.c_float
    args = [None] * 13 + [8, 8, 8, None, 0, f(0), None] + [f(1)] * 11 + [1, 0, None, None]
    rc = lib.nv_hydro_stage(*args)
    assert rc == 1001
    assert b"null" in lib.nv_last_error()


def test_rejects_bad_enums_and_sizes(lib):
    f = ctypes.c_float
    p = ctypes.c_void_p(16)
    q = ctypes.c_void_p(32)
    base = [p] * 9 + [q] * 4
    rc = lib.nv_hydro_stage(*base, 8, 8, 8, None, 0, f(0), None, *[f(1)] * 11, 7, 0, None, None)
    assert rc == 1001 and b"solver" in lib.nv_last_error()
    rc = lib.nv_hydro_stage(*base, 8, 8, 8, None, 0, f(0), None, *[f(1)] * 11, 1, 9, None, None)
    assert rc == 1001 and b"limiter" in lib.nv_last_error()
    rc = lib.nv_hydro_stage(*base, 0, 8, 8, None, 0, f(0), None, *[f(1)] * 11, 1, 0, None, None)
    assert rc == 1001 and b"grid" in lib.nv_last_error()
    # output aliasing the stencil input U is refused (stage 2 may only alias U0)
    rc = lib.nv_hydro_stage(*([p] * 13), 8, 8, 8, None, 0, f(0), None, *[f(1)] * 11, 1, 0, None, None)
    assert rc == 1001 and b"alias" in lib.nv_last_error()
    rc = lib.nv_ionize_apply(p, 4, 4, p, 1, f(0.1), f(0.1), p, p, p, 8, 8, 8, None, None, None, None, f(0), f(1),
                             None, None)
    assert rc == 1001 and b"exactly one" in lib.nv_last_error()
    rc = lib.nv_cast_rays(p, 8, 8, 8, f(0), f(0), f(0), f(1), f(0), 10, f(0.1), f(1), f(1), f(1), f(1), p, 4, 4,
                          None)
    assert rc == 1001 and b"dr" in lib.nv_last_error()


def test_error_is_raised_in_python(lib):
    with pytest.raises(_lib.NebulaHipError, match="nv_smooth_xi"):
        _lib.call("nv_smooth_xi", None, None, 4, 4, 4, None)
