"""The square root of the spiky kernel (sqrt_rn_normal in pbf_device.h)
against the build's IEEE sqrtf on every float it can receive.

The spiky gradient takes the root of r^2 in [1e-10, h^2]; sqrt_rn_normal is
claimed exact on [2^-96, FLT_MAX].  The test library checks every bit pattern
of that range on the device, which makes the claim a measured fact."""
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
    return lib


def test_spiky_sqrt_is_ieee_on_its_whole_range(gpu):
    lib = selftest_lib()
    lo, hi = 0x0F800000, 0x7F7FFFFF  # 2^-96 .. FLT_MAX
    bad, first = C.c_ulonglong(0), C.c_uint32(0)
    assert lib.fsi_selftest_sqrt(lo, hi, C.byref(bad), C.byref(first)) == 0
    assert bad.value == 0, f"{bad.value} mismatches, first at bits {first.value:#010x}"
    assert first.value == 0xFFFFFFFF
