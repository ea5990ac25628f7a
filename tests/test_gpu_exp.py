"""GPU: the FP64 exponential of the integrand and far-field kernels
(Cody-Waite reduction with a split ln 2, degree-13 Taylor polynomial) against
the host's exp, within one unit in the last place, over the arguments the
kernels pass (-kappa r <= 0) and the reduction's edge cases."""
import numpy as np
import pytest

from beep3_amd._lib import dp, lib
from beep3_amd.device import Device

pytestmark = pytest.mark.gpu


def test_exp_within_one_ulp():
    rng = np.random.default_rng(9)
    ln2 = np.log(2.0)
    x = np.concatenate([
        -rng.uniform(0, 1, 200_000),           # near field: kappa r < 1
        -rng.uniform(0, 50, 200_000),          # far field
        -np.exp(rng.uniform(-40, np.log(700), 100_000)),  # tiny to large magnitudes
        -(np.arange(1, 1000) + 0.5) * ln2,     # reduction boundaries |r| = ln2 / 2
        -np.arange(0, 1000) * ln2,             # r = 0 exactly (k ln2)
        [0.0, -0.0, -1e-300, -5e-324, -707.9, -708.0, -708.5, -745.0, -1e4],
    ])
    y = np.empty_like(x)
    d = Device(0)
    d._check(lib().beep_hip_selftest_exp(d._h, len(x), x.ctypes.data_as(dp), y.ctypes.data_as(dp)))
    d.close()
    ref = np.exp(x)
    normal = x >= -708.0
    ulp = np.abs(y[normal] - ref[normal]) / np.spacing(ref[normal])
    print("max ulp", ulp.max(), "mean", ulp.mean(), "exact", np.mean(ulp == 0))
    assert ulp.max() <= 1.0
    assert np.all(y[~normal] == 0.0)
