"""Workgroup residency on the device, as the launch shapes of the kernels
assume it: a workgroup of four or more waves holds ceil(waves / 4) wave
slots on every one of the CU's four SIMDs, so at three waves per SIMD (168
registers) one five-wave workgroup fits per CU, three four-wave ones and
four three-wave ones.  The product (`k_apply_v`) therefore runs four-wave
workgroups, the FP64 gradient at p6 three-wave ones, the element-derivative
and pseudo-Hessian kernels at p6 stay at 128 registers or fewer (two
five-wave workgroups per CU).  Measured with spinning probe workgroups
(`f25_selftest_residency`)."""
import ctypes as C

import pytest

pytestmark = pytest.mark.gpu


def _groups(nt, vgprs, lds=0):
    from fem25_amd import _lib
    g = C.c_int(-1)
    _lib.check(_lib.load().f25_selftest_residency(nt, vgprs, lds, C.byref(g)), "f25_selftest_residency")
    return g.value


@pytest.mark.parametrize("nt,vgprs,expect", [
    (320, 168, 1),   # five waves at three per SIMD: five of twelve slots
    (384, 168, 1),   # six waves
    (256, 168, 3),   # four waves: every slot
    (192, 168, 4),   # three waves: every slot
    (320, 128, 2),   # five waves at four per SIMD
    (256, 128, 4),
])
def test_workgroups_per_cu(nt, vgprs, expect):
    assert _groups(nt, vgprs) == expect


def test_lds_limits_as_expected():
    # 65,000 bytes: two per CU of the 160 KiB, whatever the registers allow
    assert _groups(192, 128, 65000) == 2
    assert _groups(256, 128, 30000) == 4


def test_argument_checks():
    from fem25_amd import _lib
    g = C.c_int(-1)
    L = _lib.load()
    assert L.f25_selftest_residency(100, 168, 0, C.byref(g)) == _lib.F25_ERR_ARG
    assert L.f25_selftest_residency(256, 100, 0, C.byref(g)) == _lib.F25_ERR_ARG
    assert L.f25_selftest_residency(256, 168, 200 * 1024, C.byref(g)) == _lib.F25_ERR_ARG
