"""GPU: the 16x16x4 MFMA lane layout the M2L kernels rely on.

A (16x4): lane l holds A[l % 16][l / 16]; B (4x16): lane l holds
B[l / 16][l % 16]; accumulator element r of lane l holds D[l / 16 + 4 r][l % 16]
for the FP64 instruction and D[4 (l / 16) + r][l % 16] for the FP32 one."""
import numpy as np
import pytest

from beep3_amd._lib import dp, lib
from beep3_amd.device import Device

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fp32", [0, 1])
def test_mfma_16x16x4_layout(fp32):
    d = Device(0)
    rng = np.random.default_rng(fp32)
    a = rng.integers(-4, 5, size=(16, 4)).astype(np.float64)
    b = rng.integers(-4, 5, size=(4, 16)).astype(np.float64)
    raw = np.zeros(256)
    d._check(lib().beep_hip_selftest_mfma(d._h, fp32, a.ctypes.data_as(dp), b.ctypes.data_as(dp),
                                          raw.ctypes.data_as(dp)))
    d.close()
    ab = a @ b
    got = np.zeros((16, 16))
    for l in range(64):
        for r in range(4):
            row = 4 * (l // 16) + r if fp32 else l // 16 + 4 * r
            got[row, l % 16] = raw[4 * l + r]
    print("fp32" if fp32 else "fp64", "raw lane 0..1:", raw[:8])
    assert np.array_equal(got, ab)
