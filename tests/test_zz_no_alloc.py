"""nv_cast_rays_noalloc (the brick-sampled ray cast with its brick array in the library's static scratch) against
nv_ray_bricks + nv_cast_rays_ex with a caller buffer: bit-identical fronts on two calls in a row, the second on a
smaller grid, for the float and the FP64 long-ray column; a grid whose bricks exceed the scratch is rejected before
any launch. Every call goes through the C ABI."""
import ctypes
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
f = ctypes.c_float


def say(*a):
    print(*a, flush=True)


def cast_args(shape, star, Q, rays):
    dx = 1.0 / max(shape)
    L = [n * dx for n in shape]
    diag = math.sqrt(sum(v * v for v in L)) + 0.6
    dr = 0.5 * dx
    return dict(star=star, Q=Q, dr=dr, nsteps=int(diag / dr) + 1, dx=dx, rcap=diag, L=L, rays=rays)


def run(L, entry, rho, shape, a, rfront, wide, bricks=None):
    torch = pytest.importorskip("torch")
    strm = torch.cuda.current_stream().cuda_stream
    common = (*shape, f(a["star"][0]), f(a["star"][1]), f(a["star"][2]), f(a["Q"]), f(a["dr"]), a["nsteps"],
              f(a["dx"]), f(a["rcap"]), f(a["L"][0]), f(a["L"][1]), f(a["L"][2]), rfront.data_ptr(), *a["rays"], wide,
              strm)
    if entry == "noalloc":
        return L.nv_cast_rays_noalloc(rho.data_ptr(), *common)
    rc = L.nv_ray_bricks(rho.data_ptr(), *shape, bricks.data_ptr(), strm)
    assert rc == 0, L.nv_last_error()
    return L.nv_cast_rays_ex(rho.data_ptr(), bricks.data_ptr(), *common)


CASES = [((96, 96, 96), (0.52, 0.47, 0.5), 0.2 ** 3 / 3, (128, 256), 1),
         ((40, 36, 44), (0.31, 0.52, 0.44), 0.12 ** 3 / 3, (64, 128), 2)]


def test_noalloc_cast_matches_caller_buffer():
    torch = pytest.importorskip("torch")
    from nebulaviewer_amd import _lib
    L = _lib.load()
    for shape, star, Q, rays, seed in CASES:
        rng = np.random.default_rng(seed)
        rho = torch.from_numpy(rng.lognormal(0.0, 0.5, shape).astype(np.float32)).cuda()
        a = cast_args(shape, star, Q, rays)
        bricks = torch.zeros(L.nv_ray_bricks_size(*shape) // 4, dtype=torch.float32, device="cuda")
        for wide in (0, 1):
            say(f"[noalloc] grid {shape} rays {rays} wide={wide}")
            got = torch.full(rays, -1.0, dtype=torch.float32, device="cuda")
            ref = torch.full(rays, -2.0, dtype=torch.float32, device="cuda")
            assert run(L, "noalloc", rho, shape, a, got, wide) == 0, L.nv_last_error()
            assert run(L, "ex", rho, shape, a, ref, wide, bricks) == 0, L.nv_last_error()
            torch.cuda.synchronize()
            g, r = got.cpu().numpy(), ref.cpu().numpy()
            assert np.array_equal(g, r), f"{shape} wide={wide}: max |d| {np.abs(g - r).max()}"
            # the fronts lie inside the box, short of the ray cap: the casts did march
            assert 0.0 < r.min() and r.max() < a["rcap"]


def test_noalloc_rejects_grids_beyond_the_scratch():
    torch = pytest.importorskip("torch")
    from nebulaviewer_amd import _lib
    L = _lib.load()
    shape = (256, 256, 256)
    rho = torch.zeros(1, dtype=torch.float32, device="cuda")       # never read: the size check comes first
    rf = torch.zeros((8, 16), dtype=torch.float32, device="cuda")
    a = cast_args(shape, (0.5, 0.5, 0.5), 1e-3, (8, 16))
    rc = run(L, "noalloc", rho, shape, a, rf, 0)
    assert rc == 1001
    assert "static scratch" in L.nv_last_error().decode()
