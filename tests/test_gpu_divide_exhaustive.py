"""The float-only exact division of nv_divide.h (the FP32 stage's divisions), checked exhaustively on the GPU
that runs it: the refined hardware reciprocal is RN(1/b) for every normal b below 2^126, and the division
equals the IEEE quotient for all 2^46 pairs of significands a, b in [1, 2). Outside [1, 2) the kernels' guard
admits only operands for which the computation is an exactly scaled copy of one of these pairs. The helper
library is tests/csrc/libnv_divcheck.so (built by __graft_entry__.build())."""
import ctypes
import os
import sys

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
    return L


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
