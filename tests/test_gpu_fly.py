"""On-the-fly near field (BEEP_NEAR_ON_THE_FLY): every mat-vec integrates the
explicit pairs instead of reading stored coefficients, the semantics of the
reference's OpenCL path (BEM_kernels_instant, src/bem/opencl_bem_kernels.cl:
237-323).  Its operator is the precomputed one to rounding, in every
precision and with either far field, and the oracle pins it directly."""
import time

import numpy as np
import pytest

import oracle
from beep3_amd import configs
from beep3_amd._lib import BeepError
from beep3_amd.device import Device

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _device(a, mode, kappa, far="direct", order=5, precs=("fp64", "fp32")):
    d = Device(0)
    d.set_far_field(far, order, 40)
    d.set_near_mode(mode)
    d.set_geometry(a)
    d.set_tree(a)
    for p in precs:
        d.precompute(kappa, p)
    return d


def test_on_the_fly_vs_stored_and_oracle(born_scene):
    """2,562 patches, direct far field.  The same pair values are summed in
    another order (row groups of 64 instead of 256), so FP64 agrees to
    rounding and FP32 to FP32 rounding; no coefficient memory is held."""
    sc, a, orc = born_scene
    n = orc.n
    x = np.random.default_rng(31).normal(size=2 * n)
    dp = _device(a, "precompute", 0.125)
    df = _device(a, "on_the_fly", 0.125)
    e = {p: rel(df.matvec(x, p), dp.matvec(x, p)) for p in ("fp64", "fp32", "fp32_resid")}
    e_orc = rel(df.matvec(x, "fp64"), orc.matvec(x))
    held = (dp.memory_held(), df.memory_held())
    nnz = int(np.sum(np.diff(a["leaf_ptr"]) * np.diff(a["nb_ptr"])))
    print("on the fly vs stored:", e, "vs oracle (fp64):", e_orc, "bytes held", held, "nnz", nnz)
    assert e["fp64"] < 1e-13
    assert e["fp32_resid"] < 1e-6 and e["fp32"] < 1e-6
    assert e_orc < 1e-12
    # FP64 (32 B) and FP32 (16 B) coefficient sets not held
    assert held[0] - held[1] >= 0.9 * 48 * nnz
    # no stored entries to poke; the self terms still are
    with pytest.raises(BeepError):
        df.debug_poke("fp64", 0, [0.0, 0.0, 0.0, 0.0])
    # the SpMV settings do not apply
    with pytest.raises(BeepError):
        df.autotune_spmv("fp64", [1, 2], reps=1)
    dp.close()
    df.close()


def test_on_the_fly_solve_matches_stored(born_scene):
    """Mixed-precision and FP64 solves give the stored-coefficient solution
    to the solver tolerance, with the same iteration counts."""
    sc, a, orc = born_scene
    n = orc.n
    b = np.random.default_rng(32).normal(size=2 * n)
    res = {}
    for mode in ("precompute", "on_the_fly"):
        d = _device(a, mode, 0.125)
        res[mode] = {m: d.solve(b, tol=1e-10, mode=m) for m in ("fp64", "mixed")}
        d.close()
    for m in ("fp64", "mixed"):
        xp, sp = res["precompute"][m]
        xf, sf = res["on_the_fly"][m]
        print(m, "solution rel diff", rel(xf, xp), "iterations", sp["iterations"], sf["iterations"])
        assert rel(xf, xp) < 1e-8
        assert abs(sp["iterations"] - sf["iterations"]) <= 1


@pytest.fixture(scope="module")
def off_centre():
    from beep3_amd.scene import Scene

    b = configs.off_centre()
    sc = Scene(b.Dsolvent, b.kappa, b.quad, b.qual, b.nbsize)
    sc.add_instance(b.library[0], 4.0)
    sc.build()
    a = sc.arrays()
    return b, a, oracle.Oracle(a, b.kappa, b.Dsolvent)


def test_on_the_fly_with_fmm_full_size(off_centre):
    """40,962 patches with the Chebyshev FMM (explicit pairs the leaf pass
    covers keep the full rule): on the fly against stored coefficients, and
    200 sampled rows against the oracle's exact far sum."""
    b, a, orc = off_centre
    n = orc.n
    x = np.random.default_rng(33).normal(size=2 * n)
    dp = _device(a, "precompute", b.kappa, far="fmm")
    df = _device(a, "on_the_fly", b.kappa, far="fmm")
    e = {p: rel(df.matvec(x, p), dp.matvec(x, p)) for p in ("fp64", "fp32", "fp32_resid")}
    ms = {}
    for p in ("fp64", "fp32"):
        df.synchronize()
        t0 = time.perf_counter()
        for _ in range(3):
            df.matvec(x, p)
        ms[p] = (time.perf_counter() - t0) / 3 * 1e3
    y = df.matvec(x, "fp64")
    rows = np.random.default_rng(34).choice(n, 200, replace=False)
    ref = orc.matvec(x, rows=rows)
    got = np.vstack([y[rows], y[n + rows]])
    e_orc = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    print("40,962 FMM: on the fly vs stored", e, "sampled rows vs oracle", e_orc, "ms per mat-vec", ms)
    assert e["fp64"] < 1e-13
    assert e["fp32_resid"] < 1e-6 and e["fp32"] < 1e-6
    assert e_orc < 1e-4
    dp.close()
    df.close()
