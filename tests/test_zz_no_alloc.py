"""GPU: the allocation-free mat-vec (beep_hip_matvec_device_noalloc) against
beep_hip_matvec_device on the same context, through the C ABI: the same bits
on two calls in a row, the second on a smaller problem (the shared scratch
array still holds the first call's contents), for every precision; an FMM
context is refused."""
import numpy as np
import pytest

from beep3_amd import configs
from beep3_amd._lib import BeepError, lib, vp
from beep3_amd.device import Device
from beep3_amd.scene import Scene

pytestmark = pytest.mark.gpu


def _scene(subdivisions):
    b = configs.born_ion(subdivisions=subdivisions)
    sc = Scene(b.Dsolvent, b.kappa, b.quad, b.qual, b.nbsize)
    sc.add_instance(b.library[0], 2.0)
    sc.build()
    return sc.arrays()


def _device(a, far="direct"):
    d = Device(0)
    d.set_far_field(far, 5, 40)
    d.set_geometry(a)
    d.set_tree(a)
    d.precompute(0.125, "fp64")
    d.precompute(0.125, "fp32")
    return d


def test_noalloc_matvec_matches_matvec_device():
    for sub in (4, 3):  # 2,562 patches, then 642
        a = _scene(sub)
        n = a["area"].shape[0]
        print(f"allocation-free mat-vec, {n} patches", flush=True)
        d = _device(a)
        x = np.random.default_rng(sub).normal(size=2 * n)
        dx, d1, d2 = d.alloc(16 * n), d.alloc(16 * n), d.alloc(16 * n)
        d.upload(dx, x)
        for p in ("fp64", "fp32", "fp32_resid"):
            d.matvec_device(dx, d1, p)
            d._check(lib().beep_hip_matvec_device_noalloc(d._h, vp(dx), vp(d2),
                                                          {"fp64": 0, "fp32": 1, "fp32_resid": 2}[p]))
            d.synchronize()
            y1 = d.download(np.empty(2 * n), d1)
            y2 = d.download(np.empty(2 * n), d2)
            assert np.all(np.isfinite(y2))
            assert np.array_equal(y1, y2), (n, p)
        for ptr in (dx, d1, d2):
            d.free(ptr)
        d.close()


def test_noalloc_refuses_the_fmm():
    a = _scene(3)
    d = _device(a, far="fmm")
    n = a["area"].shape[0]
    dx, dy = d.alloc(16 * n), d.alloc(16 * n)
    with pytest.raises(BeepError, match="direct far field"):
        d._check(lib().beep_hip_matvec_device_noalloc(d._h, vp(dx), vp(dy), 0))
    d.free(dx)
    d.free(dy)
    d.close()
