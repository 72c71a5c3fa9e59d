"""The square root and the division of the pair kernels (sqrt_rn_normal and
div_rn_normal in pbf_device.h) against the build's IEEE sqrtf and a / b.

The spiky gradient takes the root of r^2 in [1e-10, h^2]; sqrt_rn_normal is
claimed exact on [2^-96, FLT_MAX].  The test library checks every bit pattern
of that range on the device, which makes the claim a measured fact.

The pair divisions (the spiky gradient's, s_corr's w / wq) use div_rn_normal
when the host finds their operands in numerators [2^-80, 2^60], denominators
[2^-40, 2^40]; the test library compares it with a / b on every numerator of
that range, of both signs, for 64 denominators across theirs."""
import ctypes as C
import os

import pytest

from conftest import PKG

pytestmark = pytest.mark.gpu


def selftest_lib():
    path = os.path.join(PKG, "lib", "libfsi_selftest.so")
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built (make -C realtime-fsi-engine_amd)")
    lib = C.CDLL(path)
    lib.fsi_selftest_sqrt.argtypes = [C.c_uint32, C.c_uint32, C.POINTER(C.c_ulonglong), C.POINTER(C.c_uint32)]
    lib.fsi_selftest_sqrt.restype = C.c_int
    lib.fsi_selftest_div.argtypes = [C.c_uint32, C.c_uint32, C.POINTER(C.c_float), C.c_uint32,
                                     C.POINTER(C.c_ulonglong), C.POINTER(C.c_uint32)]
    lib.fsi_selftest_div.restype = C.c_int
    return lib


def test_spiky_sqrt_is_ieee_on_its_whole_range(gpu):
    lib = selftest_lib()
    lo, hi = 0x0F800000, 0x7F7FFFFF  # 2^-96 .. FLT_MAX
    bad, first = C.c_ulonglong(0), C.c_uint32(0)
    assert lib.fsi_selftest_sqrt(lo, hi, C.byref(bad), C.byref(first)) == 0
    assert bad.value == 0, f"{bad.value} mismatches, first at bits {first.value:#010x}"
    assert first.value == 0xFFFFFFFF


def test_pair_division_is_ieee_on_its_range(gpu):
    import numpy as np

    lib = selftest_lib()
    rng = np.random.default_rng(7)
    bs = [2.0 ** -40, 2.0 ** 40, 1e-5, 0.05, 0.1, 1.0, 3.0, 1234.5678]
    bs += list(2.0 ** rng.uniform(-40.0, 40.0, 56))
    b = np.asarray(bs, dtype=np.float32)
    lo, hi = (127 - 80) << 23, (127 + 60) << 23  # 2^-80 .. 2^60
    bad, first = C.c_ulonglong(0), C.c_uint32(0)
    arr = b.ctypes.data_as(C.POINTER(C.c_float))
    assert lib.fsi_selftest_div(lo, hi, arr, len(b), C.byref(bad), C.byref(first)) == 0
    assert bad.value == 0, f"{bad.value} mismatches, first numerator bits {first.value:#010x}"
    assert first.value == 0xFFFFFFFF
