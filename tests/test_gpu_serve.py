"""GPU: the solve server (include/beep_serve.h).  A request's result is
bit-identical whether it is solved alone or in a mix of concurrent requests
(continuous batching over several slots), and a Born ion solved through the
server matches the Born-Kirkwood energy."""
import time

import numpy as np
import pytest

import oracle
from beep3_amd import configs
from beep3_amd.scene import Mesh, Scene
from beep3_amd.serve import Server

pytestmark = pytest.mark.gpu

# (icosphere level, shape seed, charges, charge seed)
REQUESTS = [(3, 11, 50, 1), (4, 12, 200, 2), (5, 13, 500, 3), (4, 14, 80, 4), (3, 15, 120, 5),
            (5, 16, 300, 6)]


def _meshes():
    return [Mesh.star(lvl, 25.0, 8, 0.04, s).charges_walk(n, 3.8, 2.0, 0.4, cs)
            for lvl, s, n, cs in REQUESTS]


def _scene(m):
    return Scene(80.0, 0.125, configs.QUAD, configs.QUAL, 1500).add_instance(m, 2.0)


def _collect(srv, ids, timeout=300):
    out, t0 = {}, time.time()
    while len(out) < len(ids) and time.time() - t0 < timeout:
        r = srv.poll(timeout=2.0)
        if r is None:
            continue
        assert r["status"] == 0, r["message"]
        f, h = srv.take_fh(r["id"], r["n_patches"])
        out[ids[r["id"]]] = (r, f, h)
    assert len(out) == len(ids), "requests did not finish"
    return out


@pytest.mark.parametrize("far", ["fmm", "direct"])
def test_alone_and_in_mix_are_identical(far):
    meshes = _meshes()
    opts = dict(keep_fh=True, far_field=far, fmm_order=5, fmm_leaf=40, budget_gib=32)
    alone = {}
    srv = Server(slots=1, **opts)
    for k, m in enumerate(meshes):
        alone.update(_collect(srv, {srv.submit(_scene(m)): k}))
    srv.close()
    # the mix: every request twice, all submitted at once onto 4 slots
    srv = Server(slots=4, **opts)
    ids = {}
    for rep in range(2):
        for k, m in enumerate(meshes):
            ids[srv.submit(_scene(m))] = (k, rep)
    mix = _collect(srv, ids)
    srv.close()
    slots_used = {r["slot"] for r, _, _ in mix.values()}
    print(f"{far}: energies", [round(alone[k][0]["energy"], 6) for k in range(len(meshes))],
          "slots used in the mix", sorted(slots_used))
    assert len(slots_used) > 1
    for (k, rep), (r, f, h) in mix.items():
        ra, fa, ha = alone[k]
        assert r["converged"] == 1
        assert r["energy"] == ra["energy"], (k, rep, r["energy"], ra["energy"])
        assert np.array_equal(f, fa) and np.array_equal(h, ha), (k, rep)


def test_poses_of_one_mesh_alone_and_in_mix():
    """Poses of one shared mesh (two instances, FMM far field): every result,
    energy included, has the same bits alone and in a mix whatever request
    of the same mesh ran before it."""
    mesh = configs.pose_mesh(5)
    rng = np.random.default_rng(11)
    poses = [configs.random_pose(rng) for _ in range(5)]

    def scene(k):
        q, off = poses[k]
        sc = Scene(80.0, 0.125, configs.QUAD, configs.QUAL, 1500)
        return sc.add_instance(mesh, 2.0, offset=[0.0, 0.0, 0.0]).add_instance(
            mesh, 2.0, quat=q, offset=off)

    opts = dict(keep_fh=True, far_field="fmm", fmm_order=5, fmm_leaf=40, budget_gib=32)
    alone = {}
    for k in range(len(poses)):
        srv = Server(slots=1, **opts)  # a fresh server: nothing ran before
        alone.update(_collect(srv, {srv.submit(scene(k)): k}))
        srv.close()
    srv = Server(slots=3, **opts)
    ids = {}
    for rep in range(2):
        for k in reversed(range(len(poses))):
            ids[srv.submit(scene(k))] = (k, rep)
    mix = _collect(srv, ids)
    srv.close()
    for (k, rep), (r, f, h) in mix.items():
        ra, fa, ha = alone[k]
        assert r["converged"] == 1
        assert r["energy"] == ra["energy"], (k, rep, r["energy"], ra["energy"])
        assert np.array_equal(f, fa) and np.array_equal(h, ha), (k, rep)


def test_born_ion_through_the_server():
    m = Mesh.icosphere(4, 2.0).set_charges([[0.0, 0.0, 0.0, 1.0]])
    srv = Server(slots=2, keep_fh=False, mode="mixed")
    rid = srv.submit(_scene(m))
    r = srv.poll(timeout=120.0)
    srv.close()
    assert r is not None and r["id"] == rid and r["status"] == 0, r
    E_ref = oracle.born_energy(1.0, 2.0, 2.0, 80.0, 0.125)
    print("server Born energy", r["energy"], "analytic", E_ref)
    assert abs(r["energy"] - E_ref) / abs(E_ref) < 5e-3


def test_pose_through_server_matches_single_solve():
    """A two-instance pose solved by the server gives the same f, h bits as
    the same pose solved on its own through the BEEP object."""
    rng = np.random.default_rng(5)
    mesh = configs.pose_mesh(4)
    q, off = configs.random_pose(rng)
    b = configs.pose_pair(mesh, q, off)
    b.solve(1e-6, 100, mode="mixed")
    E_single = b.calculate_energies()
    srv = Server(slots=2, keep_fh=True, mode="mixed")
    sc = Scene(80.0, 0.125, configs.QUAD, configs.QUAL, 1500)
    sc.add_instance(mesh, 2.0, offset=[0.0, 0.0, 0.0])
    sc.add_instance(mesh, 2.0, quat=q, offset=off)
    rid = srv.submit(sc)
    out = _collect(srv, {rid: 0})
    srv.close()
    r, f, h = out[0]
    print("pose energy server", r["energy"], "single", E_single)
    assert np.array_equal(f, b.f) and np.array_equal(h, b.h)
    assert abs(r["energy"] - E_single) <= 1e-12 * abs(E_single)
    b.close()


def test_curved_mesh_through_the_server():
    """A curved (PNG1) mesh goes through the server, which has the device
    generate its point sets from the placed control points: its energy
    matches the oracle's LU solve on the host's curved arrays."""
    m = Mesh.icosphere(4, 2.0).set_charges([[0.3, 0.0, -0.2, 1.0], [-0.5, 0.4, 0.1, -0.5]]).set_curved(True)
    srv = Server(slots=2, keep_fh=False, mode="mixed")
    rid = srv.submit(Scene(80.0, 0.125, 6, 0, 1500).add_instance(m, 2.0))
    r = srv.poll(timeout=120.0)
    srv.close()
    assert r is not None and r["id"] == rid and r["status"] == 0, r
    sc = Scene(80.0, 0.125, 6, 0, 1500).add_instance(m, 2.0).build()
    o = oracle.Oracle(sc.arrays(), 0.125, 80.0)
    E_ref = o.energies(np.linalg.solve(o.assemble(), o.rhs())).sum()
    print("curved mesh through the server: energy", r["energy"], "oracle LU", E_ref)
    assert abs(r["energy"] - E_ref) / abs(E_ref) < 1e-5


def test_near_field_on_the_fly_when_coefficients_do_not_fit():
    """A budget below the stored-coefficient estimate but above the
    on-the-fly one: the auto mode integrates the near field on the fly and
    the energy matches the stored-coefficient solve."""
    m = Mesh.icosphere(4, 2.0).set_charges([[0.3, 0.0, -0.2, 1.0], [-0.5, 0.4, 0.1, -0.5]])
    probe = Server(slots=1)
    stored, fly = probe.estimate(Scene(80.0, 0.125, 6, 0, 1500).add_instance(m, 2.0))
    probe.close()
    res = {}
    for name, kw in (("stored", dict(near_mode="precompute")),
                     ("forced", dict(near_mode="on_the_fly")),
                     ("auto", dict(near_mode="auto", budget_gib=(stored + fly) / 2 / 2.0**30))):
        srv = Server(slots=1, mode="mixed", tol=1e-10, **kw)
        rid = srv.submit(Scene(80.0, 0.125, 6, 0, 1500).add_instance(m, 2.0))
        r = srv.poll(timeout=120.0)
        srv.close()
        assert r is not None and r["id"] == rid and r["status"] == 0, r
        res[name] = r
    print("near field stored / forced / auto on the fly:",
          [(k, v["energy"], v["near_on_the_fly"]) for k, v in res.items()], "estimates", stored, fly)
    assert res["stored"]["near_on_the_fly"] == 0
    assert res["forced"]["near_on_the_fly"] == 1 and res["auto"]["near_on_the_fly"] == 1
    for k in ("forced", "auto"):
        assert abs(res[k]["energy"] - res["stored"]["energy"]) <= 1e-8 * abs(res["stored"]["energy"])


def test_device_tree_and_host_tree_requests_are_identical():
    """The server's default builds each request's octree on the device after
    admission; with device_tree=False the host builds it with the scene.  The
    trees are the same, so f, h and the energy have the same bits."""
    meshes = _meshes()[:3]
    res = {}
    for dev_tree in (True, False):
        srv = Server(slots=2, keep_fh=True, far_field="fmm", fmm_order=5, device_tree=dev_tree)
        ids = {srv.submit(_scene(m)): k for k, m in enumerate(meshes)}
        res[dev_tree] = _collect(srv, ids)
        srv.close()
    for k in range(len(meshes)):
        rd, fd, hd = res[True][k]
        rh, fh, hh = res[False][k]
        print(k, "patches", rd["n_patches"], "host ms device tree / host tree", rd["host_ms"], rh["host_ms"])
        assert rd["energy"] == rh["energy"]
        assert np.array_equal(fd, fh) and np.array_equal(hd, hh)
