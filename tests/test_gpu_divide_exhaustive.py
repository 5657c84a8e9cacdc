"""The float-only exact division and square root of nv_divide.h (the FP32 stage's), checked exhaustively on
the GPU that runs them: sqrt_rn equals the IEEE square root for every float from 2^-96 up, the refined hardware
reciprocal is RN(1/b) for every normal b below 2^126, and the division equals the IEEE quotient for all 2^46
pairs of significands a, b in [1, 2). Outside [1, 2) the kernels' guard admits only operands for which the
computation is an exactly scaled copy of one of these pairs. The helper library is tests/csrc/libnv_divcheck.so
(built by __graft_entry__.build())."""
import ctypes
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def lib():
    import torch
    assert torch.cuda.is_available()
    path = os.path.join(HERE, "csrc", "libnv_divcheck.so")
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: run __graft_entry__.build() first")
    L = ctypes.CDLL(path)
    L.nv_check_recip_f.restype = ctypes.c_longlong
    L.nv_check_recip_f.argtypes = [ctypes.c_void_p]
    L.nv_check_div_f.restype = ctypes.c_longlong
    L.nv_check_div_f.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p]
    L.nv_check_sqrt_rn.restype = ctypes.c_longlong
    L.nv_check_sqrt_rn.argtypes = [ctypes.c_void_p]
    return L


def test_sqrt_is_correctly_rounded(lib):
    say("[divide] sqrt_rn over every float from 2^-96 through +inf")
    rec = (ctypes.c_uint32 * 33)()
    n = lib.nv_check_sqrt_rn(rec)
    assert n == 0, f"{n} square roots differ, first (x, s) bits: {list(rec)[1:9]}"


def test_reciprocal_is_correctly_rounded(lib):
    say("[divide] recip_f over every normal b below 2^126")
    rec = (ctypes.c_uint32 * 33)()
    n = lib.nv_check_recip_f(rec)
    assert n == 0, f"{n} reciprocals differ, first (b, y) bits: {list(rec)[1:9]}"


def test_division_all_significand_pairs(lib):
    rec = (ctypes.c_uint32 * 33)()
    total = 0
    for lo in range(0, 1 << 23, 1 << 21):           # four slices of 2^21 divisors x 2^23 numerators
        say(f"[divide] div_fn over b significands [{lo:#x}, {lo + (1 << 21):#x}) x all a")
        n = lib.nv_check_div_f(lo, lo + (1 << 21), 1, rec)
        assert n >= 0, "HIP error in the check"
        total += n
    assert total == 0, f"{total} quotients differ, first (a, b) bits: {list(rec)[1:9]}"


def edge_operands(rng, groups, G):
    """numerators and divisors on and around every edge of the guard: zeros of both signs, subnormal and tiny
    normal numerators, quotients near 2^+-64, divisors near 2^-37 and 2^91 and beyond, infinities and NaN"""
    F32 = np.float32
    pow2 = lambda e: np.ldexp(np.float64(1.0), e)
    bvals = [pow2(e) * m for e in (-126, -60, -38, -37, -36, -1, 0, 1, 30, 89, 90, 91, 92, 125) for m in (1.0, 1.5, 1.9999999)]
    bvals += [np.inf, 6.0, 98.0, 3.0, 1e-3, 0.0625]
    avals = [0.0, -0.0, np.inf, -np.inf, np.nan, 1.0, -3.0]
    avals += [s * pow2(e) * m for e in (-149, -140, -127, -126, -110, -104, -103, -102, -101, -100, -65, -64, -63, 0,
                                        40, 63, 64, 65, 100, 127) for m in (1.0, 1.25, 1.9999999) for s in (1, -1)]
    b = rng.choice(np.array(bvals, np.float64), groups).astype(F32)
    wide = rng.uniform(size=groups) < 0.3
    b[wide] = rng.lognormal(0.0, 20.0, int(wide.sum())).astype(F32)        # divisors over many binades
    a = rng.choice(np.array(avals, np.float64), groups * G).astype(F32)
    # quotients placed next to 2^+-64: a = b q for q just inside and outside the guard
    near = rng.uniform(size=groups * G) < 0.25
    qn = rng.choice(np.array([pow2(-64) * 0.9999999, pow2(-64), pow2(-64) * 1.0000001, pow2(64) * 0.9999999,
                              pow2(64), pow2(64) * 1.0000001]), groups * G).astype(F32)
    bb = np.repeat(b, G)
    with np.errstate(all="ignore"):
        a[near] = (qn[near] * bb[near]).astype(F32)
    return a, b


@pytest.mark.parametrize("G", [1, 3, 4])
def test_division_guard_edges(lib, G):
    """div_fn on the guard's edges (and far outside it) equals numpy's IEEE float32 division bit for bit, NaN
    for NaN"""
    import torch
    say(f"[divide] guard edges, groups of {G}")
    L = lib
    L.nv_div_fn_apply.restype = ctypes.c_int
    L.nv_div_fn_apply.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int, ctypes.c_int]
    rng = np.random.default_rng(40 + G)
    groups = 1 << 16
    a, b = edge_operands(rng, groups, G)
    ga, gb = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    gq = torch.empty_like(ga)
    assert L.nv_div_fn_apply(ga.data_ptr(), gb.data_ptr(), gq.data_ptr(), groups, G) == 0
    q = gq.cpu().numpy()
    with np.errstate(all="ignore"):
        want = a / np.repeat(b, G)
    same = (q.view(np.uint32) == want.view(np.uint32)) | (np.isnan(q) & np.isnan(want))
    bad = np.flatnonzero(~same)
    assert bad.size == 0, [(float(a[i]), float(np.repeat(b, G)[i]), float(q[i]), float(want[i])) for i in bad[:5]]


def test_sqrt_edges(lib):
    """sqrt_rn below 2^-96 (the IEEE path), at zero, negative and NaN arguments equals numpy's float32 sqrt"""
    import torch
    say("[divide] sqrt_rn edge arguments")
    L = lib
    L.nv_sqrt_rn_apply.restype = ctypes.c_int
    L.nv_sqrt_rn_apply.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    rng = np.random.default_rng(7)
    bits = rng.integers(0, 1 << 32, 1 << 20, dtype=np.uint64).astype(np.uint32)
    bits[: 1 << 18] &= np.uint32(0x0fffffff)            # small magnitudes, subnormals among them
    x = bits.view(np.float32).copy()
    x[:4] = [0.0, -0.0, -1.0, np.inf]
    gx = torch.from_numpy(x).cuda()
    gs = torch.empty_like(gx)
    assert L.nv_sqrt_rn_apply(gx.data_ptr(), gs.data_ptr(), x.size) == 0
    s = gs.cpu().numpy()
    with np.errstate(all="ignore"):
        want = np.sqrt(x)
    same = (s.view(np.uint32) == want.view(np.uint32)) | (np.isnan(s) & np.isnan(want))
    assert same.all(), [(float(x[i]), float(s[i]), float(want[i])) for i in np.flatnonzero(~same)[:5]]
