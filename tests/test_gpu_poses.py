"""GPU: pose batches through the C ABI.

* the shared near field (beep_hip_share_near): the mesh's own coefficient
  sets applied to both instances of a pose plus the listed inter-instance
  pairs, against the FP64 oracle given the same explicit pairs row by row;
* a pose solved in a batch has the same bits alone and among others;
* the pose energies against the reference's own split (one octree over both
  instances, every pose's pairs integrated), within the difference the two
  near/far splits make.
"""
import numpy as np
import pytest

import oracle
from beep3_amd import configs
from beep3_amd.device import Device
from beep3_amd.poses import PoseBatch
from beep3_amd.scene import Scene

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def close_pose(mesh_arrays, gap=1.5):
    """Identity rotation, instance 1 moved along x until the two surfaces are
    about `gap` apart (instance 0 at the origin)."""
    c = mesh_arrays["centroid"] - mesh_arrays["centroid"].mean(axis=0)
    return np.array([1.0, 0.0, 0.0, 0.0]), np.array([c[:, 0].max() - c[:, 0].min() + gap, 0.0, 0.0])


def pose_pairs(cent, n1, rc):
    """Pairs between the instances closer than rc, both orders."""
    from scipy.spatial import cKDTree
    t = cKDTree(cent[n1:])
    pi, pj = [], []
    for i, js in enumerate(t.query_ball_point(cent[:n1], rc)):
        for j in js:
            pi += [i, n1 + j]
            pj += [n1 + j, i]
    return np.array(pi, dtype=np.int32), np.array(pj, dtype=np.int32)


def explicit_rows(a2, a1, n1, pi, pj, rows):
    """The pose arrays with one leaf per row whose neighbourhood is exactly
    that row's explicit columns (the mesh's own near list on its copy, plus
    its inter-instance pairs); rows not sampled get none."""
    lp, li, npt, ni = a1["leaf_ptr"], a1["leaf_idx"], a1["nb_ptr"], a1["nb_idx"]
    leaf_of = np.empty(n1, dtype=np.int64)
    for L in range(len(lp) - 1):
        leaf_of[li[lp[L]:lp[L + 1]]] = L
    N = 2 * n1
    lists = [np.zeros(0, dtype=np.int32)] * N
    for r in rows:
        copy, i = divmod(int(r), n1)
        L = leaf_of[i]
        lists[r] = np.concatenate([ni[npt[L]:npt[L + 1]] + copy * n1, pj[pi == r]]).astype(np.int32)
    nb_ptr = np.zeros(N + 1, dtype=np.int32)
    nb_ptr[1:] = np.cumsum([len(x) for x in lists])
    out = dict(a2)
    out.update(leaf_ptr=np.arange(N + 1, dtype=np.int32), leaf_idx=np.arange(N, dtype=np.int32),
               nb_ptr=nb_ptr, nb_idx=np.concatenate(lists).astype(np.int32))
    return out


@pytest.fixture(scope="module")
def small_mesh():
    return configs.pose_mesh(subdivisions=4)  # 2,562 patches, the pose generator's shape


def test_shared_near_field_vs_oracle(small_mesh):
    mesh = small_mesh
    one = Scene(80.0, 0.125, configs.QUAD, configs.QUAL, 1500).add_instance(mesh, 2.0, offset=[0, 0, 0])
    one.build()
    a1 = one.arrays()
    n1 = a1["area"].shape[0]
    q, off = close_pose(a1)
    two = Scene(80.0, 0.125, configs.QUAD, configs.QUAL, 1500)
    two.add_instance(mesh, 2.0, offset=[0, 0, 0])
    two.add_instance(mesh, 2.0, quat=q, offset=off)
    two.build()
    a2 = two.arrays()
    pi, pj = pose_pairs(a2["centroid"], n1, 3.0)
    print(f"shared near field: 2 x {n1} patches, {len(pi)} inter-instance pairs", flush=True)
    assert len(pi) > 0
    src = Device(0)
    src.set_far_field("direct")
    src.set_geometry(a1)
    src.set_tree(a1)
    src.precompute(0.125, "fp64")
    src.precompute(0.125, "fp32")
    d = Device(0)
    d.set_far_field("direct")
    d.set_geometry(a2)
    d.share_near(src, [0, n1], pi, pj)
    d.precompute(0.125, "fp64")
    d.precompute(0.125, "fp32")
    x = np.random.default_rng(51).normal(size=4 * n1)
    rows = np.concatenate([np.random.default_rng(52).choice(2 * n1, 150, replace=False), np.unique(pi)[:50]])
    rows = np.unique(rows)
    orc = oracle.Oracle(explicit_rows(a2, a1, n1, pi, pj, rows), 0.125, 80.0)
    ref = orc.matvec(x, rows=rows)
    errs = {p: rel(np.vstack([y[rows], y[2 * n1 + rows]]), ref)
            for p, y in ((p, d.matvec(x, p)) for p in ("fp64", "fp32", "fp32_resid"))}
    # the same pairs found on the device (beep_hip_share_near_cutoff): the
    # same operator to the bit
    e = Device(0)
    e.set_far_field("direct")
    e.set_geometry(a2)
    n_dev = e.share_near_cutoff(src, [0, n1], 3.0)
    e.precompute(0.125, "fp64")
    e.precompute(0.125, "fp32")
    same = all(np.array_equal(d.matvec(x, p), e.matvec(x, p)) for p in ("fp64", "fp32", "fp32_resid"))
    e.close()
    d.close()
    src.close()
    print("pairs found on the device:", n_dev, "host:", len(pi), "same mat-vecs:", same)
    assert n_dev == len(pi) and same
    print("shared near field vs oracle (sampled rows):", errs)
    assert errs["fp64"] < 1e-12 and errs["fp32"] < 2e-6 and errs["fp32_resid"] < 2e-7


def test_pose_batch_alone_and_among_others(small_mesh):
    rng = np.random.default_rng(7)
    poses = [configs.random_pose(rng) for _ in range(5)]
    pb = PoseBatch(small_mesh, slots=3, keep_fh=True, far_field="fmm", fmm_order=5, fmm_leaf=40)
    print("pose batch:", pb.info(), flush=True)
    alone = pb.solve(poses[:1])
    f0, h0 = pb.take_fh(0)
    mix = pb.solve(poses[3:] + poses[:1] + poses[1:3])
    f1, h1 = pb.take_fh(2)
    pb.close()
    print("alone", alone[0]["energy"], alone[0]["iterations"], "in the mix", mix[2]["energy"])
    assert alone[0]["status"] == 0 and all(r["status"] == 0 for r in mix), [r["message"] for r in mix]
    assert all(r["converged"] == 1 for r in mix)
    assert mix[2]["energy"] == alone[0]["energy"]
    assert np.array_equal(f0, f1) and np.array_equal(h0, h1)


def test_pose_batch_vs_reference_split(small_mesh):
    """The same poses through the reference's split (one octree over both
    instances) and through the batch (the mesh's own near list per instance,
    inter-instance pairs by distance): the energies differ only by the pairs
    the two splits treat differently (72-point rule vs 1 point)."""
    rng = np.random.default_rng(8)
    poses = [configs.random_pose(rng) for _ in range(3)]
    a1 = Scene(80.0, 0.125, configs.QUAD, configs.QUAL, 1500).add_instance(small_mesh, 2.0).build().arrays()
    poses.append(close_pose(a1, 2.0))
    pb = PoseBatch(small_mesh, slots=2, mode="fp64", far_field="direct", tol=1e-8)
    got = pb.solve(poses)
    pb.close()
    out = []
    for (q, off), r in zip(poses, got):
        b = configs.pose_pair(small_mesh, q, off, far_field="direct")
        b.init(precisions=["fp64"])
        b.solve(1e-8, 100, mode="fp64")
        E = b.calculate_energies()
        b.close()
        out.append((r["energy"], E, abs(r["energy"] - E) / abs(E), r["n_pairs"]))
    print("batch vs reference split (energy, energy, rel, inter pairs):", out)
    for eb, er, d, npairs in out:
        assert d < 2e-4


def test_curved_pose_batch_vs_reference_split():
    """A curved (PNG1) mesh as a pose batch: the device generates both
    instances' curved point sets from the placed control points, and the
    energies agree with the reference's split solved as one two-instance
    scene (host-side BEEP driver) to the bound of the planar comparison."""
    mesh = configs.pose_mesh(subdivisions=4).set_curved(True)
    rng = np.random.default_rng(9)
    poses = [configs.random_pose(rng) for _ in range(2)]
    pb = PoseBatch(mesh, slots=2, mode="fp64", far_field="direct", tol=1e-8)
    got = pb.solve(poses)
    pb.close()
    for (q, off), r in zip(poses, got):
        assert r["status"] == 0, r["message"]
        b = configs.pose_pair(mesh, q, off, far_field="direct")
        b.init(precisions=["fp64"])
        b.solve(1e-8, 100, mode="fp64")
        E = b.calculate_energies()
        b.close()
        d = abs(r["energy"] - E) / abs(E)
        print("curved pose: batch", r["energy"], "reference split", E, "rel", d)
        assert d < 2e-4
