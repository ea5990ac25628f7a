"""Host-side scene construction: meshes, node patches, point sets, poses and
the explicit-pair octree.  CPU only.

The node-patch point sets are checked against an independent numpy
restatement of BasicTriangle::get_constant_basis_quad_points
(src/bem/triangle.h:388-479): for every triangle around a vertex, the two
sub-triangles (centroid, vertex, edge midpoint) of the barycentric dual,
each carrying one sixth of the triangle's area."""
import numpy as np
import pytest

from beep3_amd import Mesh, Scene

R6 = (np.array([0.659027622374092, 0.659027622374092, 0.231933368553031, 0.231933368553031,
                0.109039009072877, 0.109039009072877]),
      np.array([0.231933368553031, 0.109039009072877, 0.659027622374092, 0.109039009072877,
                0.659027622374092, 0.231933368553031]),
      np.full(6, 1 / 6))
R4 = (np.array([1 / 3, 0.2, 0.6, 0.2]), np.array([1 / 3, 0.6, 0.2, 0.2]),
      np.array([-0.5625, 0.520833333333, 0.520833333333, 0.520833333333]))


def dual_points(xyz, tri, v, rule):
    """Points of the node patch of vertex v: list of (p, n, w)."""
    ru, rv, rw = rule
    out = []
    for t in np.nonzero(np.any(tri == v, axis=1))[0]:
        a, b, c = xyz[tri[t]]
        cr = np.cross(b - a, c - a)
        area = 0.5 * np.linalg.norm(cr)
        nrm = cr / np.linalg.norm(cr)
        corner = a if tri[t][0] == v else (b if tri[t][1] == v else c)
        others = [p for k, p in enumerate((a, b, c)) if tri[t][k] != v]
        cen = (a + b + c) / 3
        m1 = (corner + others[0]) / 2
        m2 = (corner + others[1]) / 2
        for P1, P2, P3 in ((cen, corner, m1), (m2, corner, cen)):
            for u, vv, w in zip(ru, rv, rw):
                out.append((P1 * (1 - u - vv) + P2 * u + P3 * vv, nrm, w * area / 6))
    return out


@pytest.mark.parametrize("sub", [0, 1, 2, 4])
def test_icosphere_counts(sub):
    m = Mesh.icosphere(sub, 1.5)
    i = m.info()
    assert i["n_vertices"] == 10 * 4**sub + 2
    assert i["n_triangles"] == 20 * 4**sub
    xyz, tri = m.vertices()
    assert np.allclose(np.linalg.norm(xyz, axis=1), 1.5)
    # outward orientation
    cr = np.cross(xyz[tri[:, 1]] - xyz[tri[:, 0]], xyz[tri[:, 2]] - xyz[tri[:, 0]])
    assert np.all(np.sum(cr * xyz[tri].mean(axis=1), axis=1) > 0)


def test_patch_points_match_numpy_restatement():
    m = Mesh.icosphere(1, 2.0).set_charges([[0, 0, 0, 1.0]])
    xyz, tri = m.vertices()
    sc = Scene(80.0, 0.1, 6, 0, 1500).add_instance(m, 2.0).build()
    a = sc.arrays()
    for v in range(xyz.shape[0]):
        mine = a["quad"][a["quad_ptr"][v]:a["quad_ptr"][v + 1]]
        ref = dual_points(xyz, tri, v, R6)
        assert mine.shape[0] == len(ref)
        # same multiset of points (order follows the triangle order)
        ref_arr = np.array([np.concatenate([p, n, [w]]) for p, n, w in ref])
        key = lambda arr: np.lexsort(np.round(arr[:, :3], 9).T)
        assert np.allclose(mine[key(mine)], ref_arr[key(ref_arr)], atol=1e-12)
        gal = a["gal"][a["gal_ptr"][v]:a["gal_ptr"][v + 1]]
        gref = dual_points(xyz, tri, v, R4)
        ga = np.array([np.concatenate([p, n, [w]]) for p, n, w in gref])
        assert np.allclose(gal[key(gal)], ga[key(ga)], atol=1e-12)
        # patch area, centroid and normal from the 4-point galerkin set
        w = ga[:, 6]
        assert a["area"][v] == pytest.approx(w.sum(), rel=1e-12)
        assert np.allclose(a["centroid"][v], (ga[:, :3] * w[:, None]).sum(0) / w.sum(), atol=1e-12)
        an = (ga[:, 3:6] * w[:, None]).sum(0)
        assert np.allclose(a["normal"][v], an / np.linalg.norm(an), atol=1e-12)


def test_point_weights_and_areas():
    m = Mesh.icosphere(3, 3.0).set_charges([[0, 0, 0, 1.0]])
    xyz, tri = m.vertices()
    total = 0.5 * np.linalg.norm(np.cross(xyz[tri[:, 1]] - xyz[tri[:, 0]],
                                          xyz[tri[:, 2]] - xyz[tri[:, 0]]), axis=1).sum()
    sc = Scene(80.0, 0.1, 6, 4, 1500).add_instance(m, 2.0).build()
    a = sc.arrays()
    # the 4-point rule's weights sum to 1 - 1e-12 (rounded 0.520833333333)
    assert a["area"].sum() == pytest.approx(total, rel=2e-12)
    qw = np.add.reduceat(a["quad"][:, 6], a["quad_ptr"][:-1])
    assert np.allclose(qw, a["area"], rtol=1e-12)
    lw = np.add.reduceat(a["qual"][:, 6], a["qual_ptr"][:-1])
    assert np.allclose(lw, 1.0, rtol=1e-12)
    val = np.bincount(tri.ravel(), minlength=xyz.shape[0])
    assert np.array_equal(np.diff(a["quad_ptr"]), 2 * val * 6)
    assert np.array_equal(np.diff(a["qual_ptr"]), 2 * val * 4)
    assert np.array_equal(np.diff(a["gal_ptr"]), 2 * val * 4)


def test_centroid_collocation_single_points():
    m = Mesh.icosphere(2, 1.0).set_charges([[0, 0, 0, 1.0]])
    a = Scene(80.0, 0.1, 0, 0, 1500).add_instance(m, 2.0).build().arrays()
    assert np.array_equal(np.diff(a["qual_ptr"]), np.ones(a["area"].size))
    assert np.allclose(a["qual"][:, :3], a["centroid"]) and np.allclose(a["qual"][:, 6], 1.0)
    assert np.allclose(a["quad"][:, 6], a["area"])


def test_rigid_pose():
    m = Mesh.icosphere(2, 2.0).set_charges([[0.1, 0.2, -0.3, 1.0], [-0.5, 0.1, 0.2, -0.5]])
    q = np.array([0.3, -0.5, 0.2, 0.7])
    q /= np.linalg.norm(q)
    off = np.array([30.0, -4.0, 11.0])
    a0 = Scene(80.0, 0.1, 6, 0, 1500).add_instance(m, 2.0).build().arrays()
    a1 = Scene(80.0, 0.1, 6, 0, 1500).add_instance(m, 2.0, quat=q, offset=off).build().arrays()
    d0 = np.linalg.norm(a0["centroid"][:, None] - a0["charges"][None, :, :3], axis=2)
    d1 = np.linalg.norm(a1["centroid"][:, None] - a1["charges"][None, :, :3], axis=2)
    assert np.allclose(d0, d1, atol=1e-12)
    assert np.allclose(a1["charges"][:, :3].mean(0), off, atol=1e-12)
    assert np.allclose(np.sum(a0["normal"] * a0["normal"], 1), 1.0)
    dots0 = np.sum(a0["normal"][:, None] * a0["normal"][None], axis=2)
    dots1 = np.sum(a1["normal"][:, None] * a1["normal"][None], axis=2)
    assert np.allclose(dots0, dots1, atol=1e-12)


@pytest.mark.parametrize("quat, offset, expect", [
    ([1, 0, 0, 0], [0, 0, 0], [5, 7, 9]),    # null rotation and translation
    ([0, 1, 0, 0], [0, 0, 0], [5, -7, -9]),  # half turn about x
    ([0, 1, 0, 0], [1, 2, 3], [6, -5, -6]),  # then the centre moved to (1, 2, 3)
])
def test_pose_known_answers(quat, offset, expect):
    """The known answers of tools/test.py:1-32 (Vector::change_coordinate_frame
    with the quaternions (1,0,0,0) and (0,1,0,0)), applied to a charge at
    (5, 7, 9) of a mesh whose charge centre is the origin."""
    m = Mesh.icosphere(1, 1.0).set_charges([[5, 7, 9, 1.0], [-5, -7, -9, 1.0]])
    a = Scene(80.0, 0.1, 0, 0, 1500).add_instance(m, 2.0, quat=quat, offset=offset).build().arrays()
    assert np.allclose(a["charges"][0, :3], expect, atol=1e-12)
    assert np.allclose(a["charges"][1, :3], 2 * np.asarray(offset) - np.asarray(expect), atol=1e-12)


def test_star_mesh_and_random_walk_charges():
    m = Mesh.star(3, 25.0, 8, 0.04, 3).charges_walk(300, 3.8, 2.0, 0.4, 3)
    xyz, tri = m.vertices()
    r = np.linalg.norm(xyz, axis=1)
    assert 15.0 < r.min() < r.max() < 35.0 and r.std() > 0.1
    a = Scene(80.0, 0.125, 6, 0, 1500).add_instance(m, 2.0).build().arrays()
    q = a["charges"]
    steps = np.linalg.norm(np.diff(q[:, :3], axis=0), axis=1)
    assert np.mean(np.isclose(steps, 3.8)) > 0.9
    # every charge lies inside the surface by the vertex radius along its direction
    d = q[:, :3] / np.maximum(np.linalg.norm(q[:, :3], axis=1, keepdims=True), 1e-12)
    near = np.argmax(d @ (xyz / r[:, None]).T, axis=1)
    assert np.all(np.linalg.norm(q[:, :3], axis=1) < r[near])


@pytest.mark.parametrize("nbsize", [60, 300, 1500])
def test_octree_explicit_pairs(nbsize):
    m = Mesh.icosphere(4, 2.0).set_charges([[0, 0, 0, 1.0]])
    sc = Scene(80.0, 0.1, 6, 0, nbsize).add_instance(m, 2.0).build()
    a = sc.arrays()
    N = a["area"].size
    lp, li, npt, ni = a["leaf_ptr"], a["leaf_idx"], a["nb_ptr"], a["nb_idx"]
    assert np.array_equal(np.sort(li), np.arange(N))  # leaves partition the patches
    pairs = 0
    for L in range(len(lp) - 1):
        rows = li[lp[L]:lp[L + 1]]
        cols = ni[npt[L]:npt[L + 1]]
        assert np.unique(cols).size == cols.size  # neighbourhood boxes are disjoint
        assert np.all(np.isin(rows, cols))        # a leaf is its own neighbour
        pairs += rows.size * cols.size
        # explicit columns are spatially close: within 2 leaf-box diagonals
        d = np.linalg.norm(a["centroid"][rows][:, None] - a["centroid"][cols][None], axis=2)
        assert d.max() < 4.0 * 2 * 1.7321 * 2.0 / 1.0
    assert pairs == sc.sizes().n_near_pairs
    # optimize_neighbourhoods target: mean explicit interactions per patch <= nbsize
    assert pairs / N <= nbsize + 1


def test_bad_arguments_raise():
    from beep3_amd._lib import BeepError
    with pytest.raises(BeepError, match="rule"):
        Scene(80.0, 0.1, 5, 0, 1500)
    with pytest.raises(BeepError):
        Scene(80.0, 0.1, 6, 0, 1500).build()


def test_off_round_trip(tmp_path):
    m = Mesh.icosphere(2, 3.0).set_charges([[0.1, 0, 0, 1.0]])
    p = tmp_path / "ico.off"
    m.write_off(str(p))
    q = tmp_path / "ico.xyzqr"
    q.write_text("# x y z q r\n0.1 0 0 1.0 1.5\n")
    m2 = Mesh.read(str(p), charges=str(q))
    a1 = Scene(80.0, 0.1, 6, 0, 1500).add_instance(m, 2.0).build().arrays()
    a2 = Scene(80.0, 0.1, 6, 0, 1500).add_instance(m2, 2.0).build().arrays()
    for k in ("centroid", "normal", "area", "quad", "charges"):
        assert np.allclose(a1[k], a2[k], atol=1e-13), k


def test_gts_reader(tmp_path):
    """GTS stores faces as edge triples; the reader recovers the vertex
    triples (shared vertex of consecutive edges) and so the same patches."""
    m = Mesh.icosphere(1, 2.0).set_charges([[0, 0, 0, 1.0]])
    xyz, tri = m.vertices()
    edges, eidx, faces = [], {}, []
    for t in tri:
        fe = []
        for a, b in ((t[0], t[1]), (t[1], t[2]), (t[2], t[0])):
            k = (min(a, b), max(a, b))
            if k not in eidx:
                eidx[k] = len(edges)
                edges.append(k)
            fe.append(eidx[k] + 1)
        faces.append(fe)
    lines = [f"{len(xyz)} {len(edges)} {len(faces)}"]
    lines += [f"{x:.17g} {y:.17g} {z:.17g}" for x, y, z in xyz]
    lines += [f"{a + 1} {b + 1}" for a, b in edges]
    lines += [f"{a} {b} {c}" for a, b, c in faces]
    p = tmp_path / "ico.gts"
    p.write_text("# icosphere\n" + "\n".join(lines) + "\n")
    m2 = Mesh.read(str(p)).set_charges([[0, 0, 0, 1.0]])
    a1 = Scene(80.0, 0.1, 6, 0, 1500).add_instance(m, 2.0).build().arrays()
    a2 = Scene(80.0, 0.1, 6, 0, 1500).add_instance(m2, 2.0).build().arrays()
    assert np.allclose(a1["area"], a2["area"], atol=1e-13)
    assert np.allclose(a1["normal"], a2["normal"], atol=1e-13)


def test_bad_mesh_file(tmp_path):
    from beep3_amd._lib import BeepError
    p = tmp_path / "bad.off"
    p.write_text("OFF\n3 1 0\n0 0 0\n1 0 0\n0 1 0\n4 0 1 2 0\n")
    with pytest.raises(BeepError, match="triangles"):
        Mesh.read(str(p))
    ply = tmp_path / "x.ply"
    ply.write_text("ply\n")
    with pytest.raises(BeepError, match="off or .gts"):
        Mesh.read(str(ply))


def test_surface_export_matches_the_host_point_sets():
    """The surface handed to the device (placed vertices, triangles, the
    triangles around each vertex with its corner) regenerates, through the
    numpy restatement, the point sets the host builds for a posed two-instance
    scene; a scene left to device points exports the same surface."""
    m = Mesh.star(2, 25.0, 8, 0.04, 3).charges_walk(20, 3.8, 2.0, 0.4, 3)
    q = np.array([0.8, 0.2, -0.3, 0.4])
    scenes = []
    for host in (True, False):
        sc = Scene(80.0, 0.125, 6, 0, 1500, host_points=host)
        sc.add_instance(m, 2.0, offset=[0.0, 0.0, 0.0]).add_instance(m, 2.0, quat=q,
                                                                   offset=[55.0, 3.0, -2.0])
        scenes.append(sc.build())
    a, surf = scenes[0].arrays(), scenes[0].surface()
    surf2 = scenes[1].surface()
    for k in surf:
        assert np.array_equal(surf[k], surf2[k]), k
    N = a["area"].shape[0]
    xyz, tri = surf["verts"], surf["tris"]
    assert xyz.shape == (N, 3) and surf["vt_ptr"][-1] == surf["vt_tri"].shape[0]
    for v in range(N):
        ts = surf["vt_tri"][surf["vt_ptr"][v]:surf["vt_ptr"][v + 1]]
        cs = surf["vt_corner"][surf["vt_ptr"][v]:surf["vt_ptr"][v + 1]]
        assert np.all(tri[ts, cs] == v)
        assert sorted(ts) == sorted(np.nonzero(np.any(tri == v, axis=1))[0])
    for v in range(0, N, 7):
        mine = a["quad"][a["quad_ptr"][v]:a["quad_ptr"][v + 1]]
        ref = np.array([np.concatenate([p, n, [w]]) for p, n, w in dual_points(xyz, tri, v, R6)])
        key = lambda arr: np.lexsort(np.round(arr[:, :3], 8).T)
        assert np.allclose(mine[key(mine)], ref[key(ref)], atol=1e-11)
    # the device-point scene exports the pointer arrays but no points
    b = scenes[1].arrays()
    assert np.array_equal(b["quad_ptr"], a["quad_ptr"]) and b["quad"].shape[0] == 0
