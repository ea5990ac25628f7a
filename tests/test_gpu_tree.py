"""The octree and near list built on the device (beep_hip_build_tree) against
the host's (beep_scene_tree, itself pinned to the oracle's plain restatement
of octree.h in tests/test_octree.py): the same leaves, leaf contents and
neighbourhood contents in the same order, so the mat-vec over it has the same
bits."""
import time

import numpy as np
import pytest

from beep3_amd import configs
from beep3_amd.device import Device
from beep3_amd.scene import Scene

pytestmark = pytest.mark.gpu

KEYS = ("leaf_ptr", "leaf_idx", "nb_ptr", "nb_idx")


def _compare(sc: Scene, nbsize: int, matvec: bool):
    a, s = sc.arrays(), sc.sizes()
    d = Device(0)
    d.set_far_field("direct")
    d.set_geometry(a)
    t0 = time.perf_counter()
    nl, nn = d.build_tree(list(s.cube_centre), s.cube_edge, nbsize)
    ms = (time.perf_counter() - t0) * 1e3
    got = d.tree_arrays()
    print(f"{a['area'].shape[0]} patches: {nl} leaves, {nn} neighbourhood entries, device build {ms:.1f} ms")
    for k in KEYS:
        assert np.array_equal(got[k], a[k]), k
    if matvec:
        x = np.random.default_rng(3).normal(size=2 * a["area"].shape[0])
        h = Device(0)
        h.set_far_field("direct")
        h.set_geometry(a)
        h.set_tree(a)
        for dev in (d, h):
            dev.precompute(sc.kappa, "fp64")
        assert np.array_equal(d.matvec(x, "fp64"), h.matvec(x, "fp64"))
        h.close()
    d.close()


@pytest.mark.parametrize("subdiv,nbsize", [(4, 1500), (5, 300), (4, 40)])
def test_device_tree_single_molecule(subdiv, nbsize):
    m = configs.globular_mesh(subdiv, 3, 200, 3)
    sc = Scene(80.0, 0.125, configs.QUAD, configs.QUAL, nbsize).add_instance(m, 2.0).build()
    _compare(sc, nbsize, matvec=subdiv == 4)


def test_device_tree_one_leaf_and_uneven_instances():
    """642 patches under nbsize 1,500: the root is the only leaf (level 0).
    Then three molecules of uneven sizes side by side (2,562 + 642 +
    10,242 patches, one rotated)."""
    small = configs.globular_mesh(3, 3, 50, 4)
    sc = Scene(80.0, 0.125, configs.QUAD, configs.QUAL, 1500).add_instance(small, 2.0).build()
    _compare(sc, 1500, matvec=True)
    mid, big = configs.globular_mesh(4, 5, 100, 5), configs.globular_mesh(5, 6, 300, 6)
    sc = Scene(80.0, 0.125, configs.QUAD, configs.QUAL, 600)
    sc.add_instance(mid, 2.0, offset=[0.0, 0.0, 0.0])
    sc.add_instance(small, 2.0, quat=[0.6, -0.2, 0.7, 0.3], offset=[48.0, 5.0, -3.0])
    sc.add_instance(big, 2.0, offset=[-20.0, 55.0, 10.0])
    _compare(sc.build(), 600, matvec=False)


def test_device_tree_poses():
    m = configs.pose_mesh(5)
    rng = np.random.default_rng(5)
    for k in range(2):
        q, off = configs.random_pose(rng, 30.0, 60.0 if k else 35.0)
        sc = Scene(80.0, 0.125, configs.QUAD, configs.QUAL, 1500)
        sc.add_instance(m, 2.0, quat=None, offset=[0.0, 0.0, 0.0])
        sc.add_instance(m, 2.0, quat=q, offset=off)
        _compare(sc.build(), 1500, matvec=False)


def test_device_tree_globular_full_size():
    """163,842 patches (the mat-vec benchmark's surface); then the steady
    state of both paths: the device build against the host octree inside the
    scene build plus its upload (beep_hip_set_tree)."""
    b = configs.globular()
    sc = Scene(b.Dsolvent, b.kappa, b.quad, b.qual, b.nbsize).add_instance(b.library[0], 2.0).build()
    _compare(sc, b.nbsize, matvec=False)
    a, s = sc.arrays(), sc.sizes()
    d = Device(0)
    d.set_far_field("direct")
    d.set_geometry(a)
    dev_ms, host_ms, set_ms = [], [], []
    for rep in range(6):
        d.synchronize()
        t0 = time.perf_counter()
        d.build_tree(list(s.cube_centre), s.cube_edge, b.nbsize)
        d.synchronize()
        dev_ms.append((time.perf_counter() - t0) * 1e3)
    for rep in range(4):
        built = {}
        for tree in (False, True):
            sc2 = Scene(b.Dsolvent, b.kappa, b.quad, b.qual, b.nbsize, host_points=False, tree=tree)
            t0 = time.perf_counter()
            sc2.add_instance(b.library[0], 2.0).build()
            built[tree] = (time.perf_counter() - t0, sc2)
        host_ms.append((built[True][0] - built[False][0]) * 1e3)
        d.synchronize()
        t0 = time.perf_counter()
        d.set_tree(built[True][1].arrays())
        d.synchronize()
        set_ms.append((time.perf_counter() - t0) * 1e3)
    d.close()
    print(f"steady state, ms: device build {np.median(dev_ms[1:]):.2f} (calls {np.round(dev_ms, 2)}); "
          f"host octree {np.median(host_ms):.2f} + set_tree {np.median(set_ms):.2f}")
