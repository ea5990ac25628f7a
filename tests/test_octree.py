"""The product's host octree (beep3_amd/csrc/host_octree.cpp, built for
speed) against the oracle's plain restatement of src/fmm/octree.h under the
BEM policy of src/beep/beep.cpp:1585-1641 (oracle/octree_ref.cpp): leaves,
their contents and their neighbourhood contents must be identical, in order,
on single molecules, two-instance poses and small neighbourhood targets."""
import numpy as np
import pytest

import oracle
from beep3_amd import configs
from beep3_amd.scene import Mesh, Scene


def _check(sc: Scene, nbsize: int):
    sc.build()
    a = sc.arrays()
    s = sc.sizes()
    ref = oracle.octree_csr(a["centroid"], list(s.cube_centre), s.cube_edge, nbsize)
    for k in ("leaf_ptr", "leaf_idx", "nb_ptr", "nb_idx"):
        assert np.array_equal(a[k], ref[k]), k
    return int(a["leaf_ptr"].shape[0] - 1)


@pytest.mark.parametrize("subdiv,nbsize", [(4, 1500), (5, 300), (6, 1500), (6, 120)])
def test_octree_single_molecule(subdiv, nbsize):
    m = configs.globular_mesh(subdiv, 3, 200, 3)
    sc = Scene(80.0, 0.125, configs.QUAD, configs.QUAL, nbsize)
    sc.add_instance(m, 2.0)
    leaves = _check(sc, nbsize)
    assert leaves > 1


def test_octree_poses():
    m = configs.pose_mesh(5)
    rng = np.random.default_rng(5)
    for k in range(3):
        q, off = configs.random_pose(rng, 30.0, 60.0 if k else 35.0)
        sc = Scene(80.0, 0.125, configs.QUAD, configs.QUAL, 1500)
        sc.add_instance(m, 2.0, quat=None, offset=[0.0, 0.0, 0.0])
        sc.add_instance(m, 2.0, quat=q, offset=off)
        _check(sc, 1500)


def test_octree_globular_full_size():
    """163,842 patches (the mat-vec benchmark's surface)."""
    b = configs.globular()
    sc = Scene(b.Dsolvent, b.kappa, b.quad, b.qual, b.nbsize)
    sc.add_instance(b.library[0], 2.0)
    assert _check(sc, b.nbsize) > 1000
