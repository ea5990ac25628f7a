"""GPU parity at the full sizes of the measured workloads, through the C ABI:

* globular surface, 163,842 patches, 2,000 random-walk charges: sampled rows
  of the FP64, FP32 and FP32-coefficient/FP64-sum mat-vecs against the FP64
  oracle (direct and FMM far field), and the mixed-precision solve's energy,
  f and h against the plain FP64 solve with the exact direct far field;
* two instances of the 40,962-patch surface (a_lm seed 7) in the first
  seeded rigid poses: sampled rows against the oracle, and the server's
  result equal to the single solve's;
* the serving mix (2,562 .. 163,842 patches): every request's result is the
  same alone and among the others.
"""
import time

import numpy as np
import pytest

import oracle
from beep3_amd import configs
from beep3_amd.device import Device
from beep3_amd.scene import Scene
from beep3_amd.serve import Server

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def sampled_rows(orc, y, rows):
    return np.vstack([y[rows], y[orc.n + rows]])


def _device(a, kappa, far, order=5, leaf=40, precs=("fp64", "fp32")):
    d = Device(0)
    d.set_far_field(far, order, leaf)
    d.set_geometry(a)
    d.set_tree(a)
    for p in precs:
        d.precompute(kappa, p)
    return d


# ---------------- globular surface, 163,842 patches ----------------

@pytest.fixture(scope="module")
def globular_case():
    b = configs.globular()
    sc = Scene(b.Dsolvent, b.kappa, b.quad, b.qual, b.nbsize)
    sc.add_instance(b.library[0], 2.0)
    sc.build()
    a = sc.arrays()
    assert a["area"].shape[0] == 163842
    return b, a, oracle.Oracle(a, b.kappa, b.Dsolvent)


# bounds, a few times the measured errors: the direct far field leaves only
# rounding (FP64 2.0e-15; FP32 coefficients and sums 8.7e-7; FP32
# coefficients with FP64 sums 6.3e-9); the order-5 FMM adds its truncation
# error (1.75e-5)
GLOBULAR_BOUNDS = {
    ("direct", "fp64"): 1e-13, ("direct", "fp32"): 3e-6, ("direct", "fp32_resid"): 3e-8,
    ("fmm", "fp64"): 5e-5, ("fmm", "fp32"): 5e-5, ("fmm", "fp32_resid"): 5e-5,
}


@pytest.mark.parametrize("far", ["direct", "fmm"])
def test_globular_rows(globular_case, far):
    b, a, orc = globular_case
    print(f"globular 163,842 patches, {far} far field", flush=True)
    d = _device(a, b.kappa, far)
    x = np.random.default_rng(31).normal(size=2 * orc.n)
    rows = np.random.default_rng(32).choice(orc.n, 200, replace=False)
    ref = orc.matvec(x, rows=rows)
    errs = {}
    for p in ("fp64", "fp32", "fp32_resid"):
        errs[p] = rel(sampled_rows(orc, d.matvec(x, p), rows), ref)
    d.close()
    print(f"{far}: sampled-row rel err vs oracle", errs)
    for p, e in errs.items():
        assert e < GLOBULAR_BOUNDS[(far, p)], (far, p, e)


def test_globular_mixed_solve_vs_fp64_direct():
    """Mixed precision (FP32 mat-vecs, FP64 refinement, order-5 FMM) against
    the plain FP64 solve with the exact direct far field: energy within
    1e-5 relative, f and h L2-relative errors stated."""
    print("globular 163,842 patches: FP64 direct solve", flush=True)
    ref = configs.globular(far_field="direct")
    ref.init(precisions=["fp64"])
    ref.solve(1e-6, 100, mode="fp64")
    E_ref = ref.calculate_energies()
    f_ref, h_ref = ref.f.copy(), ref.h.copy()
    it_ref = ref.run_info["iterations"]
    ref.close()
    print("globular 163,842 patches: mixed FMM solve", flush=True)
    mix = configs.globular(far_field="fmm", fmm_order=5, fmm_leaf=40)
    mix.init(precisions=["fp32"])
    mix.solve(1e-6, 100, mode="mixed")
    E = mix.calculate_energies()
    info = dict(mix.run_info)
    ef, eh = rel(mix.f, f_ref), rel(mix.h, h_ref)
    mix.close()
    eE = abs(E - E_ref) / abs(E_ref)
    print(f"energy {E:.8f} vs {E_ref:.8f} kJ/mol: rel {eE:.2e}; f L2 {ef:.2e}, h L2 {eh:.2e};",
          f"iterations {it_ref} (FP64) / {info['iterations']} (mixed),",
          f"refinements {info['refinements']}")
    assert info["converged"] == 1
    assert eE <= 1e-5
    # the FMM truncation dominates: measured about 7e-5 (f) and 2e-4 (h)
    assert ef < 5e-4 and eh < 1e-3


# ---------------- two instances of the 40,962-patch surface, seeded poses ----------------

@pytest.fixture(scope="module")
def pose_case():
    mesh = configs.pose_mesh()  # 40,962 patches, a_lm seed 7, 2,000 charges
    rng = np.random.default_rng(7)  # the pose generator of configs.pose_scenes
    poses = [configs.random_pose(rng) for _ in range(3)]
    return mesh, poses


def _pose_scene(mesh, q, off, host_points=True):
    sc = Scene(80.0, 0.125, configs.QUAD, configs.QUAL, 1500, host_points=host_points)
    sc.add_instance(mesh, 2.0, offset=[0.0, 0.0, 0.0])
    sc.add_instance(mesh, 2.0, quat=q, offset=off)
    return sc


def test_pose_rows(pose_case):
    mesh, poses = pose_case
    q, off = poses[0]
    sc = _pose_scene(mesh, q, off).build()
    a = sc.arrays()
    assert a["area"].shape[0] == 2 * 40962
    orc = oracle.Oracle(a, 0.125, 80.0)
    x = np.random.default_rng(41).normal(size=2 * orc.n)
    rows = np.random.default_rng(42).choice(orc.n, 200, replace=False)
    ref = orc.matvec(x, rows=rows)
    out = {}
    # measured: direct 2.1e-15 (FP64), 6.1e-7 (FP32); order-5 FMM 3.0e-5
    for far, bounds in (("direct", {"fp64": 1e-13, "fp32": 2e-6}),
                        ("fmm", {"fp64": 1e-4, "fp32": 1e-4})):
        print(f"pose 0 (81,924 patches), {far} far field", flush=True)
        d = _device(a, 0.125, far)
        for p, tol in bounds.items():
            out[(far, p)] = e = rel(sampled_rows(orc, d.matvec(x, p), rows), ref)
            assert e < tol, (far, p, e)
        d.close()
    print("pose sampled-row rel err vs oracle", out)


def _collect(srv, ids, timeout=600):
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


def test_poses_server_equals_single_solve(pose_case):
    """Each pose solved through the server (all three at once on two slots)
    gives the f and h bits of the same pose solved on its own."""
    mesh, poses = pose_case
    single = []
    for k, (q, off) in enumerate(poses):
        print(f"pose {k}: single solve", flush=True)
        b = configs.pose_pair(mesh, q, off, fmm_order=5, fmm_leaf=40)
        b.solve(1e-6, 100, mode="mixed")
        single.append((b.f.copy(), b.h.copy(), b.calculate_energies(), b.run_info["iterations"]))
        b.close()
    print("poses through the server", flush=True)
    srv = Server(slots=2, keep_fh=True, mode="mixed", fmm_order=5, fmm_leaf=40)
    ids = {srv.submit(_pose_scene(mesh, q, off, host_points=False)): k
           for k, (q, off) in enumerate(poses)}
    out = _collect(srv, ids)
    srv.close()
    for k in range(len(poses)):
        r, f, h = out[k]
        fs, hs, Es, its = single[k]
        print(f"pose {k}: energy server {r['energy']:.10f} single {Es:.10f}, iterations {its}")
        assert r["converged"] == 1
        assert np.array_equal(f, fs) and np.array_equal(h, hs), k
        assert abs(r["energy"] - Es) <= 1e-12 * abs(Es)


def test_pose_batch_full_size(pose_case):
    """The first seeded poses through the pose batch (the 40,962-patch
    mesh's own near field computed once, inter-instance pairs per pose):
    each pose's result is the same alone and among the others, and its
    energy is close to the reference split's (one octree over both
    instances), the difference being the pairs the two splits treat
    differently."""
    from beep3_amd.poses import PoseBatch

    mesh, poses = pose_case
    pb = PoseBatch(mesh, slots=3, keep_fh=True, mode="mixed", far_field="fmm", fmm_order=5, fmm_leaf=40)
    print("pose batch, 2 x 40,962 patches:", pb.info(), flush=True)
    alone = pb.solve(poses[:1])
    f0, h0 = pb.take_fh(0)
    allp = pb.solve(poses)
    f1, h1 = pb.take_fh(0)
    pb.close()
    assert all(r["status"] == 0 and r["converged"] == 1 for r in alone + allp), [r["message"] for r in allp]
    assert allp[0]["energy"] == alone[0]["energy"]
    assert np.array_equal(f0, f1) and np.array_equal(h0, h1)
    for k, (q, off) in enumerate(poses):
        b = configs.pose_pair(mesh, q, off, fmm_order=5, fmm_leaf=40)
        b.solve(1e-6, 100, mode="mixed")
        E = b.calculate_energies()
        b.close()
        d = abs(allp[k]["energy"] - E) / abs(E)
        print(f"pose {k}: batch {allp[k]['energy']:.6f} reference split {E:.6f} rel {d:.2e}, "
              f"{allp[k]['n_pairs']} inter pairs, {allp[k]['iterations']} iterations")
        assert d < 2e-4


# ---------------- serving mix, 2,562 .. 163,842 patches ----------------

def test_serving_mix_full_sizes_alone_and_in_mix():
    """Requests of the serving stream's generator, one of each size up to
    163,842 patches plus the next arrivals: each result has the same bits
    solved alone (one slot, one request at a time) and in a mix (every
    request twice, all submitted at once onto four slots)."""
    meshes = configs.serving_meshes(shapes=2)
    sched = configs.serving_schedule(rate=50.0, duration=6.0, seed=11)
    picked, seen = [], set()
    for r in sched:  # the first arrival of each size
        if r["level"] not in seen:
            seen.add(r["level"])
            picked.append(r)
    picked += [r for r in sched[:12] if r not in picked][:3]
    assert {r["level"] for r in picked} == {4, 5, 6, 7}
    print("serving requests (level, charges):", [(r["level"], r["n_charges"]) for r in picked],
          flush=True)
    opts = dict(keep_fh=True, mode="mixed", budget_gib=96)
    alone = {}
    srv = Server(slots=1, **opts)
    for k, req in enumerate(picked):
        sc, _m = configs.serving_request(meshes, req)
        alone.update(_collect(srv, {srv.submit(sc): k}))
    srv.close()
    print("alone done; now the mix", flush=True)
    srv = Server(slots=4, **opts)
    ids = {}
    keep = []
    for rep in range(2):
        for k, req in enumerate(picked):
            sc, m = configs.serving_request(meshes, req)
            keep.append(m)
            ids[srv.submit(sc)] = (k, rep)
    mix = _collect(srv, ids)
    srv.close()
    for (k, rep), (r, f, h) in mix.items():
        ra, fa, ha = alone[k]
        assert r["converged"] == 1, (k, rep)
        assert r["n_patches"] == ra["n_patches"]
        assert r["energy"] == ra["energy"], (k, rep, r["energy"], ra["energy"])
        assert np.array_equal(f, fa) and np.array_equal(h, ha), (k, rep)
    print("sizes:", sorted({r["n_patches"] for r, _, _ in mix.values()}),
          "slots used:", sorted({r["slot"] for r, _, _ in mix.values()}))
