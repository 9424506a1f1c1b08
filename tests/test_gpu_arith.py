"""Device checks of the in-range square root and division of the neighbour passes.

npx_sqrt_in / npx_div_in (novaphy_amd/csrc/npx_device.hpp) are the compiler's
correctly rounded sqrtf and '/' sequences without the steps that act only on
inputs outside a stated range.  These tests compare them bit for bit with
sqrtf and '/' on the device: the square root on every float from 2^-96 to the
largest finite one, the division on seeded operand pairs inside the stated
range and on the operands the PBF passes form for h in [2^-6, 2^6].
"""
import ctypes as C
import os

import pytest

pytestmark = pytest.mark.gpu

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hip", "libnpx_arith_check.so")


@pytest.fixture(scope="module")
def lib():
    if not os.path.isfile(LIB):
        pytest.fail(f"{LIB} is not built (python -m novaphy_amd.build)")
    L = C.CDLL(LIB)
    L.npxt_sqrt_check.argtypes = [C.c_uint, C.c_uint, C.POINTER(C.c_ulonglong), C.POINTER(C.c_uint)]
    L.npxt_sqrt_check.restype = C.c_int
    L.npxt_div_check.argtypes = [C.c_int, C.c_ulonglong, C.c_ulonglong, C.POINTER(C.c_ulonglong)]
    L.npxt_div_check.restype = C.c_int
    return L


def test_sqrt_in_equals_sqrtf_on_every_float_from_2_pow_minus_96(lib):
    bad, first = C.c_ulonglong(0), C.c_uint(0)
    lo, hi = 0x0F800000, 0x7F7FFFFF  # 2^-96 .. FLT_MAX
    print(f"sqrt check over bit patterns {lo:#x}..{hi:#x}", flush=True)
    assert lib.npxt_sqrt_check(lo, hi, C.byref(bad), C.byref(first)) == 0
    assert bad.value == 0, f"{bad.value} mismatches, first at {first.value:#010x}"


@pytest.mark.parametrize("mode,name", [(0, "stated range"), (1, "spiky_grad operands"), (2, "s_corr operands")])
def test_div_in_equals_division(lib, mode, name):
    out = (C.c_ulonglong * 3)()
    n = 1 << 28
    print(f"division check, mode {mode} ({name}), {n} pairs", flush=True)
    assert lib.npxt_div_check(mode, n, 0x5EED0000 + mode, out) == 0
    bad, outside, tested = out[0], out[1], out[2]
    assert tested > n // 4, f"only {tested} pairs compared"
    assert outside == 0, f"{outside} operand pairs outside the stated range"
    assert bad == 0, f"{bad} of {tested} quotients differ"
