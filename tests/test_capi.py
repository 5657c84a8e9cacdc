"""The C ABI library: it loads, exports every entry point include/nebula_hip.h declares with the argument count
the ctypes binding uses, and rejects bad arguments with a message before touching the GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

from nebulaviewer_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "nebula_hip.h")


def declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    out = {}
    for m in re.finditer(r"\b(?:int|long long|const char\*)\s+(nv_\w+)\s*\(([^)]*)\)\s*;", src):
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
    assert lib.nv_abi_version() == 5


def struct_fields(name):
    """(type, name) of each member of a typedef struct in the header, arrays expanded as name[i]"""
    src = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    body = re.search(r"typedef struct %s \{(.*?)\} %s;" % (name, name), src, re.S).group(1)
    out = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        m = re.match(r"(const\s+)?(\w+)(\s*\*)?\s*(.*)", decl)
        base = m.group(2) + ("*" if m.group(3) or "*" in m.group(4) else "")
        for item in m.group(4).split(","):
            item = item.strip().lstrip("*").strip()
            arr = re.match(r"(\w+)\[(\d+)\]", item)
            if arr:
                out += [(base, f"{arr.group(1)}[{i}]") for i in range(int(arr.group(2)))]
            else:
                out.append((base, item))
    return out


@pytest.mark.parametrize("name", ["NvStageDesc", "NvRayDesc", "NvIonDesc", "NvPackParams"])
def test_descriptor_structs_match_header(name):
    """the ctypes mirrors have the header's members in order, with pointer/int/float kinds and C sizes"""
    cls = getattr(_lib, name)
    hdr = struct_fields(name)
    py = []
    for fname, ftype in cls._fields_:
        if issubclass(ftype, ctypes.Array):
            py += [(ftype._type_, f"{fname}[{i}]") for i in range(ftype._length_)]
        else:
            py.append((ftype, fname))
    assert [n for _, n in py] == [n for _, n in hdr]
    kinds = {"float": ctypes.c_float, "int": ctypes.c_int}
    for (pt, pn), (ht, hn) in zip(py, hdr):
        if ht.endswith("*"):
            assert pt in (ctypes.c_void_p,), (pn, pt)
        else:
            assert pt is kinds[ht], (pn, ht, pt)


def test_descriptor_layout_matches_c(tmp_path):
    """sizeof and every offsetof of the descriptor structs as gcc lays out the header, against the ctypes mirrors"""
    import subprocess
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-o", str(exe), os.path.join(ROOT, "tests", "csrc", "struct_layout.c")],
                   check=True)
    c = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                     check=True).stdout.splitlines())
    for name in ("NvStageDesc", "NvRayDesc", "NvIonDesc", "NvPackParams"):
        cls = getattr(_lib, name)
        assert ctypes.sizeof(cls) == int(c[name]), name
        for fname, _ in cls._fields_:
            assert getattr(cls, fname).offset == int(c[f"{name}.{fname}"]), (name, fname)


def test_batch_plans_validate(lib):
    """plans refuse empty batches, short buffers and bad members before any launch"""
    assert lib.nv_hydro_stage_batch_plan_size(0) == -1
    need = lib.nv_hydro_stage_batch_plan_size(2)
    assert need > 0
    members = (_lib.NvStageDesc * 2)()
    buf = ctypes.create_string_buffer(int(need))
    assert lib.nv_hydro_stage_batch_plan(members, 2, 0, 1, 0, 0, buf, 8) == 1001
    assert b"plan needs" in lib.nv_last_error()
    assert lib.nv_hydro_stage_batch_plan(members, 2, 0, 1, 0, 0, buf, need) == 1001
    assert b"member 0" in lib.nv_last_error()
    rays = (_lib.NvRayDesc * 1)()
    rbuf = ctypes.create_string_buffer(int(lib.nv_cast_rays_batch_plan_size(1)))
    assert lib.nv_cast_rays_batch_plan(rays, 1, rbuf, len(rbuf)) == 1001
    # a ray job with a brick array: aligned, not overlapping its density field
    r = rays[0]
    r.rho, r.rfront = 4096, 1 << 20
    r.nx = r.ny = r.nz = 8
    r.nu, r.nphi, r.dr, r.dx = 4, 4, 0.1, 0.1
    r.bricks = 4096 + 1024
    assert lib.nv_cast_rays_batch_plan(rays, 1, rbuf, len(rbuf)) == 1001 and b"overlap" in lib.nv_last_error()
    r.bricks = (1 << 16) + 8
    assert lib.nv_cast_rays_batch_plan(rays, 1, rbuf, len(rbuf)) == 1001 and b"aligned" in lib.nv_last_error()
    r.bricks = 1 << 16
    assert lib.nv_cast_rays_batch_plan(rays, 1, rbuf, len(rbuf)) == 0
    ions = (_lib.NvIonDesc * 1)()
    ibuf = ctypes.create_string_buffer(int(lib.nv_ionize_apply_batch_plan_size(1)))
    assert lib.nv_ionize_apply_batch_plan(ions, 1, 0, ibuf, len(ibuf)) == 1001
    assert lib.nv_ionize_apply_batch_plan(ions, 1, 4, ibuf, len(ibuf)) == 1001 and b"form" in lib.nv_last_error()
    assert lib.nv_hydro_stage_batch(buf, buf, None) == 1001     # a refused plan never got its header
    assert lib.nv_hydro_stage_batch_queue_ex(buf, buf, 0, None) == 1001
    assert lib.nv_hydro_stage_batch_queue_ex(buf, buf, -1, None) == 1001 and b"wgs_per_cu" in lib.nv_last_error()


def _ion_plan_cuts(lib, specs):
    """build an ionize batch plan on the host for members (n, nstars, culled) and return (start, xoff, xcd_run)"""
    n = len(specs)
    ions = (_lib.NvIonDesc * n)()
    for d, (g, k, culled) in zip(ions, specs):
        d.rfronts, d.star_xyz, d.xi, d.dfront_h = 1 << 20, 1 << 21, 1 << 22, 1 << 23   # never dereferenced here
        d.nu, d.nphi, d.nstars, d.dx, d.width = 8, 16, k, 1.0 / g, 1.2 / g
        d.nx = d.ny = d.nz = g
        d.rf_ranges = (1 << 24) if culled else None
    size = int(lib.nv_ionize_apply_batch_plan_size(n))
    buf = ctypes.create_string_buffer(size)
    assert lib.nv_ionize_apply_batch_plan(ions, n, 2, buf, size) == 0, lib.nv_last_error()
    ints = np.frombuffer(buf.raw, np.int32)
    xoff = ints[-9:]
    start = ints[-9 - (n + 1):-9]
    return start, xoff, int(ints[9])


def test_ionize_batch_plan_xcd_cuts(lib):
    """the batch ionize pass gives XCD x the contiguous tile run [xoff[x], xoff[x+1]): the cuts cover every tile once,
    in order, and split the work (tiles weighted by the member's star count) evenly"""
    rng = np.random.default_rng(5)
    specs = [(int(rng.choice([48, 64, 96, 128, 160])), int(rng.integers(1, 6)), True) for _ in range(64)]
    for case in (specs, specs[:1], [(512, 1, False)], [(64, 3, False), (48, 1, False), (160, 4, True)]):
        start, xoff, run = _ion_plan_cuts(lib, case)
        assert start[0] == 0 and xoff[0] == 0 and xoff[8] == start[-1]
        assert np.all(np.diff(xoff) >= 0) and run == np.diff(xoff).max()
        w = np.zeros(start[-1])
        for s, (g, k, culled) in enumerate(case):
            w[start[s]:start[s + 1]] = 1.0 if k <= 1 else 1.0 + (0.25 if culled else 0.8) * (k - 1)
        per = np.array([w[xoff[x]:xoff[x + 1]].sum() for x in range(8)])
        assert per.max() - per.mean() <= w.max() + 1e-9, (per, w.max())


def test_ionize_batch_plan_chunk_depth(lib):
    """a batch member's ionize tiles march at most 16 x-planes (32 z x 16 y cells each), the single-grid pass up to
    64: the plan's tile prefix counts ceil(n/32) ceil(n/16) ceil(n/ci) tiles per member with
    ci = min(16, max(4, n jk_total / 2048)), jk_total the batch's (z, y) tile columns"""
    rng = np.random.default_rng(6)
    for specs in ([(int(rng.choice([48, 64, 96, 128, 160])), 1, False) for _ in range(64)], [(512, 1, False)],
                  [(256, 4, True)], [(24, 2, True), (40, 1, False)]):
        start, _, _ = _ion_plan_cuts(lib, specs)
        jk = [-(-g // 32) * -(-g // 16) for g, _, _ in specs]
        want = []
        for (g, _, _), c in zip(specs, jk):
            ci = min(16, max(4, g * sum(jk) // 2048))
            want.append(c * -(-g // ci))
        np.testing.assert_array_equal(np.diff(start), want)


def test_stage_batch_plan_march_depth(lib):
    """with the library's march depth, a batch member's stage tiles march at most 32 x-planes (a single grid's up to
    64): the plan's tile prefix counts ceil(n/32) ceil(n/16) ceil(n/ci) tiles of 32 z x 16 y per member with
    ci = min(32, max(4, n jk_total / 2048)); a depth in the tile code (NV_TILE_DEPTH) is taken as given"""
    rng = np.random.default_rng(7)
    for sizes, depth in (([int(rng.choice([48, 64, 96, 128, 160])) for _ in range(64)], 0), ([512], 0),
                         ([40, 24], 0), ([96, 160, 48], 16)):
        n = len(sizes)
        ds = (_lib.NvStageDesc * n)()
        for d, g in zip(ds, sizes):
            for c in range(4):
                d.u[c], d.u0[c], d.out[c] = (1 << 20) + c, (1 << 21) + c, (1 << 22) + c   # never dereferenced here
            d.xi = 1 << 23
            d.nx = d.ny = d.nz = g
            d.dt, d.dx, d.a, d.b, d.floor_, d.cn2, d.ci2, d.vmax = 1e-4, 1.0 / g, 0.0, 1.0, 1e-3, 1.0, 1.0, 6.0
        size = int(lib.nv_hydro_stage_batch_plan_size(n))
        buf = ctypes.create_string_buffer(size)
        assert lib.nv_hydro_stage_batch_plan(ds, n, 1, 1, 0, 2 | (depth << 8), buf, size) == 0, lib.nv_last_error()
        ints = np.frombuffer(buf.raw, np.int32)
        start = ints[-(16 + n + 1):-16]
        jk = [-(-g // 32) * -(-g // 16) for g in sizes]
        want = []
        for g, c in zip(sizes, jk):
            ci = min(depth, g) if depth else min(32, max(4, g * sum(jk) // 2048))
            want.append(c * -(-g // ci))
        assert start[0] == 0 and ints[7] == start[-1]
        np.testing.assert_array_equal(np.diff(start), want)


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
    # the descriptor form: a null descriptor, a form outside the NV_ION_* bits, the same argument checks
    assert lib.nv_ionize_apply_ex(None, 0, None) == 1001 and b"null" in lib.nv_last_error()
    d = _lib.NvIonDesc()
    d.rfronts, d.star_xyz, d.xi, d.dfront_h = 16, 16, 16, 16
    d.nu = d.nphi = 4
    d.nstars, d.dx, d.width = 1, 0.1, 0.1
    d.nx = d.ny = d.nz = 8
    assert lib.nv_ionize_apply_ex(ctypes.byref(d), 4, None) == 1001 and b"form" in lib.nv_last_error()
    d.dfront_f = 16
    assert lib.nv_ionize_apply_ex(ctypes.byref(d), _lib.NV_ION_FAST, None) == 1001
    assert b"exactly one" in lib.nv_last_error()
    rc = lib.nv_cast_rays(p, 8, 8, 8, f(0), f(0), f(0), f(1), f(0), 10, f(0.1), f(1), f(1), f(1), f(1), p, 4, 4,
                          None)
    assert rc == 1001 and b"dr" in lib.nv_last_error()
    # the brick array: 16-byte aligned, not overlapping the field
    rc = lib.nv_ray_bricks(ctypes.c_void_p(4096), 8, 8, 8, ctypes.c_void_p(4096 + 8), None)
    assert rc == 1001 and b"aligned" in lib.nv_last_error()
    rc = lib.nv_ray_bricks(ctypes.c_void_p(4096), 8, 8, 8, ctypes.c_void_p(4096 + 16 * 64), None)
    assert rc == 1001 and b"overlaps" in lib.nv_last_error()
    assert lib.nv_ray_bricks_size(8, 8, 8) == 16 * 10 * 9 * 9 and lib.nv_ray_bricks_size(0, 8, 8) == -1
    rc = lib.nv_cast_rays_ex(p, ctypes.c_void_p(24), 8, 8, 8, f(0), f(0), f(0), f(1), f(0.1), 10, f(0.1), f(1), f(1),
                             f(1), f(1), p, 4, 4, 0, None)
    assert rc == 1001 and b"aligned" in lib.nv_last_error()


def test_error_is_raised_in_python(lib):
    with pytest.raises(_lib.NebulaHipError, match="nv_smooth_xi"):
        _lib.call("nv_smooth_xi", None, None, 4, 4, 4, None)


def test_ray_dirs_validates(lib):
    """nv_ray_dirs refuses a missing or misaligned table and empty grids before any launch"""
    assert lib.nv_ray_dirs(8, 8, None, None) == 1001
    assert lib.nv_ray_dirs(8, 8, ctypes.c_void_p(0x1004), None) == 1001
    assert b"16-byte aligned" in lib.nv_last_error()
    assert lib.nv_ray_dirs(0, 8, ctypes.c_void_p(0x1000), None) == 1001
    assert b"bad sizes" in lib.nv_last_error()


def test_ray_range_floats(lib):
    """the front-range buffer of a ray job: the (least, greatest) pair and one pair per 256-ray workgroup"""
    assert lib.nv_ray_range_floats(512, 1024) == 2 + 2 * (512 * 1024 // 256)
    assert lib.nv_ray_range_floats(7, 9) == 2 + 2 * 1   # one 8 x 8 patch, one workgroup
    assert lib.nv_ray_range_floats(0, 9) == -1
