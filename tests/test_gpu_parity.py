"""GPU parity of the HIP mat-vec, RHS, energy coefficients and solves against
the FP64 CPU oracle (oracle/oracle.c), through the C ABI."""
import numpy as np
import pytest

import oracle
from beep3_amd import configs
from beep3_amd.device import Device

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def dev_born(born_scene):
    sc, a, orc = born_scene
    d = Device(0)
    d.set_geometry(a)
    d.set_tree(a)
    d.precompute(0.125, "fp64")
    d.precompute(0.125, "fp32")
    yield d, a, orc
    d.close()


def test_diag_and_near_only_fp64(dev_born):
    d, a, orc = dev_born
    rng = np.random.default_rng(11)
    x = rng.normal(size=2 * orc.n)
    y_dev = d.matvec(x, "fp64")
    y_orc = orc.matvec(x)
    print("fp64 matvec rel err", rel(y_dev, y_orc))
    assert rel(y_dev, y_orc) < 1e-12


# measured 1.2e-7 (FP32 coefficients and sums) and 7.3e-9 (FP32 coefficients,
# FP64 sums): the bounds are a few times that
@pytest.mark.parametrize("prec,tol", [("fp32", 5e-7), ("fp32_resid", 3e-8)])
def test_matvec_fp32(dev_born, prec, tol):
    d, a, orc = dev_born
    rng = np.random.default_rng(12)
    x = rng.normal(size=2 * orc.n)
    e = rel(d.matvec(x, prec), orc.matvec(x))
    print(prec, "matvec rel err", e)
    assert e < tol


def test_near_singular_pairs_in_fp64(born_scene):
    """The FP32 coefficient set integrates its near-singular pairs (centroids
    within about two mesh edges) in FP64 and rounds them.  Column j of the
    FP32-coefficient / FP64-sum operator (the mat-vec of e_j) against the
    oracle's column, on the rows near-singular to j: the kernel differences
    G0 - Gk of A and D cancel there (D's by about 6 / (kappa r)^2), which
    FP32 integration does not survive and FP64 does."""
    sc, a, orc = born_scene
    n = orc.n
    c, ar = a["centroid"], a["area"]
    cols = np.random.default_rng(13).choice(n, 8, replace=False)
    out = {}
    for on in (False, True):
        d = Device(0)
        d.set_near_singular_fp64(on)
        d.set_geometry(a)
        d.set_tree(a)
        d.precompute(0.125, "fp32")
        listed, over = d.near_singular_count()
        worst = 0.0
        for j in cols:
            for comp in (0, 1):  # f_j = 1 (B, D), then h_j = 1 (A, C)
                x = np.zeros(2 * n)
                x[comp * n + j] = 1.0
                y = d.matvec(x, "fp32_resid")
                ref = orc.matvec(x)
                d2 = np.sum((c - c[j]) ** 2, axis=1)
                ns = np.nonzero((d2 < 2.0 * (ar + ar[j])) & (np.arange(n) != j))[0]
                for rows in (ns, n + ns):
                    worst = max(worst, rel(y[rows], ref[rows]))
        d.close()
        out[on] = (worst, listed, over)
    print("near-singular entries: worst rel err FP32 pass", out[False][0], "FP64 pass", out[True][0],
          "; listed", out[True][1], "for", n, "rows, overflow", out[True][2])
    assert out[False][1] == 0 and out[False][2] == 0
    assert out[True][2] == 0 and 6 * n <= out[True][1] <= 24 * n
    assert out[True][0] < 1e-6
    assert out[False][0] > 10 * out[True][0]


def test_coefficient_guard_repairs_bad_entries(born_scene):
    """BEEP::sanity_check on the device, with the bad values forced from the
    host: a NaN in the FP64 set, an out-of-range value in the FP32 set and an
    infinite self term are found and recomputed; the mat-vecs return to
    their values before the damage."""
    sc, a, orc = born_scene
    nnz = int(np.sum(np.diff(a["leaf_ptr"]) * np.diff(a["nb_ptr"])))
    d = Device(0)
    d.set_geometry(a)
    d.set_tree(a)
    d.precompute(0.125, "fp64")
    d.precompute(0.125, "fp32")
    st = d.guard_stats()
    assert st["fp64"] == st["fp32"] == st["diag"] == 0, st
    x = np.random.default_rng(14).normal(size=2 * orc.n)
    y64, y32 = d.matvec(x, "fp64"), d.matvec(x, "fp32")
    d.debug_poke("fp64", nnz // 3, [np.nan, 0.0, 0.0, 0.0])
    d.debug_poke("fp32", (2 * nnz) // 3, [0.0, 1e30, 0.0, 0.0])
    d.debug_poke("diag", 5, [0.0, 0.0, np.inf, 0.0])
    assert not np.all(np.isfinite(d.matvec(x, "fp64")))
    assert rel(d.matvec(x, "fp32"), y32) > 1.0
    assert d.sanity_check() == 3
    assert d.guard_stats()["sanity_check"] == 3
    e64, e32 = rel(d.matvec(x, "fp64"), y64), rel(d.matvec(x, "fp32"), y32)
    assert d.sanity_check() == 0
    d.close()
    print("after repair: fp64 rel change", e64, "fp32", e32)
    # the FP64 entry and the self term are recomputed by the same formulas;
    # the FP32 entry is recomputed in FP64 and rounded
    assert e64 < 1e-14 and e32 < 1e-6


def test_rhs_and_energy_coeffs(dev_born):
    d, a, orc = dev_born
    b_dev = d.rhs(a["charges"], 80.0)
    assert rel(b_dev, orc.rhs()) < 1e-13
    ef, eh = d.energy_coefficients(a["inst_patch_ptr"], a["inst_charge_ptr"], a["charges"])
    # the oracle's f coefficients are Gauss-renormalised (mesh.cpp:891-925);
    # apply the same per-instance renormalisation to the device's raw ones
    ecf, ech = orc.energy_coeffs()
    ef = renormalise(ef, a)
    print("energy coefficients rel err: f", rel(ef, ecf), "h", rel(eh, ech))
    assert rel(eh, ech) < 1e-13
    assert rel(ef, ecf) < 1e-13


def renormalise(ef, a):
    """Mesh::renormalise_energy_precalcs: spread the gap between the net
    charge and -sum(ef) over the instance's patches by area."""
    ef = ef.copy()
    pp, cp = a["inst_patch_ptr"], a["inst_charge_ptr"]
    for k in range(len(pp) - 1):
        s = slice(int(pp[k]), int(pp[k + 1]))
        net = float(np.sum(a["charges"][int(cp[k]):int(cp[k + 1]), 3]))
        ef[s] -= (a["area"][s] / np.sum(a["area"][s])) * (net + float(np.sum(ef[s])))
    return ef


def test_solve_born_ion():
    b = configs.born_ion()
    b.init(precisions=["fp64", "fp32"])
    K = None
    for mode, tol_rel in (("fp64", 1e-5), ("mixed", 1e-5)):
        b.solve(1e-6, 100, mode=mode)
        E = b.calculate_energies()
        born = oracle.born_energy(1.0, 2.0, 2.0, 80.0, 0.125)
        print(mode, "energy", E, "born", born, b.run_info)
        assert abs(E - born) / abs(born) < 1e-2
        if K is None:
            import oracle as O
            o = O.Oracle(b.host_scene().arrays(), b.kappa, b.Dsolvent)
            K = o.assemble()
            x_ref = np.linalg.solve(K, o.rhs())
            E_ref = o.energies(x_ref).sum()
        print(mode, "vs oracle LU", abs(E - E_ref) / abs(E_ref))
        assert abs(E - E_ref) / abs(E_ref) < tol_rel
    b.close()


def test_cgs2_solves(dev_born, born_scene):
    """Two Gram-Schmidt passes per Arnoldi step (CGS2): the solution agrees
    with the single-pass (reference) solve and with the oracle's LU solve,
    and the Krylov basis is orthonormal to rounding."""
    d, a, orc = dev_born
    b = orc.rhs()
    x_lu = np.linalg.solve(orc.assemble(), b)
    out = {}
    for passes in (1, 2):
        d.set_gram_schmidt(passes)
        x, st = d.solve(b, tol=1e-10, max_iter=100, mode="fp64")
        out[passes] = (x, st["iterations"])
    d.set_gram_schmidt(1)
    e1, e2 = rel(out[1][0], x_lu), rel(out[2][0], x_lu)
    print("CGS vs LU", e1, "CGS2 vs LU", e2, "iterations", out[1][1], out[2][1])
    assert e1 < 1e-8 and e2 < 1e-8
    assert out[2][1] <= out[1][1] + 1


@pytest.fixture(scope="module")
def off_centre_case():
    """Config 2 at full size (40,962 patches, 200 charges)."""
    from beep3_amd.scene import Scene

    b = configs.off_centre()
    sc = Scene(b.Dsolvent, b.kappa, b.quad, b.qual, b.nbsize)
    sc.add_instance(b.library[0], 4.0)
    sc.build()
    a = sc.arrays()
    return b, a, oracle.Oracle(a, b.kappa, b.Dsolvent)


@pytest.mark.parametrize("far,order,tol", [("direct", 5, 1e-12), ("fmm", 5, 1e-4)])
def test_off_centre_rows(off_centre_case, far, order, tol):
    """Sampled rows of the full 40,962-patch FP64 mat-vec against the oracle
    (explicit pairs + exact 1-point far sum + diagonal)."""
    b, a, orc = off_centre_case
    d = Device(0)
    d.set_far_field(far, order, 40)
    d.set_geometry(a)
    d.set_tree(a)
    d.precompute(b.kappa, "fp64")
    x = np.random.default_rng(21).normal(size=2 * orc.n)
    y = d.matvec(x, "fp64")
    d.close()
    rows = np.random.default_rng(22).choice(orc.n, 200, replace=False)
    ref = orc.matvec(x, rows=rows)
    got = np.vstack([y[rows], y[orc.n + rows]])
    e = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    print(far, "sampled-row rel err", e)
    assert e < tol


def test_off_centre_energy_vs_kirkwood(off_centre_case):
    """Full-size solve (FMM far field, mixed precision) against Kirkwood's
    series for the same 200 charges in a salted sphere; the gap is the BEM
    discretisation of a 40,962-vertex sphere."""
    b, a, orc = off_centre_case
    b.init(precisions=["fp64", "fp32"])
    b.solve(1e-6, 100, mode="mixed")
    E = b.calculate_energies()
    q = a["charges"]
    ref = oracle.kirkwood_energy(q, 20.0, 4.0, 80.0, 0.1)
    print("E", E, "Kirkwood", ref, "rel", abs(E - ref) / abs(ref))
    assert abs(E - ref) / abs(ref) < 5e-3
    b.close()


def test_device_generated_points_match_host_points():
    """Point sets generated on the device from the surface give the operator
    of the host-built point sets (a posed two-instance scene): FP64 mat-vecs
    agree to rounding, FP32 to FP32 rounding, energy coefficients to rounding."""
    from beep3_amd import Mesh
    from beep3_amd.scene import Scene

    m = Mesh.star(4, 25.0, 8, 0.04, 5).charges_walk(100, 3.8, 2.0, 0.4, 5)
    sc = Scene(80.0, 0.125, 6, 0, 1500)
    sc.add_instance(m, 2.0, offset=[0.0, 0.0, 0.0])
    sc.add_instance(m, 2.0, quat=[0.6, -0.2, 0.7, 0.3], offset=[52.0, -4.0, 6.0])
    sc.build()
    a, surf = sc.arrays(), sc.surface()
    devs = []
    for mode in ("host", "device"):
        d = Device(0)
        d.set_far_field("direct")
        if mode == "host":
            d.set_geometry(a)
        else:
            d.set_surface(a, surf, 0, 6)
        d.set_tree(a)
        d.precompute(0.125, "fp64")
        d.precompute(0.125, "fp32")
        devs.append(d)
    x = np.random.default_rng(11).normal(size=2 * a["area"].shape[0])
    e64 = rel(devs[1].matvec(x, "fp64"), devs[0].matvec(x, "fp64"))
    e32 = rel(devs[1].matvec(x, "fp32"), devs[0].matvec(x, "fp32"))
    ef = [d.energy_coefficients(a["inst_patch_ptr"], a["inst_charge_ptr"], a["charges"]) for d in devs]
    ee = max(rel(ef[1][k], ef[0][k]) for k in range(2))
    print(f"device vs host points: fp64 {e64:.2e}, fp32 {e32:.2e}, energy coefficients {ee:.2e}")
    assert e64 < 1e-13 and e32 < 1e-6 and ee < 1e-13
    for d in devs:
        d.close()


def test_curved_patches_on_the_device():
    """Curved (PNG1) node patches travel as host point sets: the FP64 and
    FP32 mat-vecs match the oracle on the same arrays, and the solve's energy
    matches the oracle's LU solve."""
    from beep3_amd import Mesh
    from beep3_amd.scene import Scene

    m = Mesh.icosphere(4, 2.0).set_charges([[0.3, 0.0, -0.2, 1.0], [-0.5, 0.4, 0.1, -0.5]]).set_curved(True)
    sc = Scene(80.0, 0.125, 6, 0, 1500).add_instance(m, 2.0).build()
    a = sc.arrays()
    orc = oracle.Oracle(a, 0.125, 80.0)
    d = Device(0)
    d.set_geometry(a)
    d.set_tree(a)
    d.precompute(0.125, "fp64")
    d.precompute(0.125, "fp32")
    x = np.random.default_rng(41).normal(size=2 * orc.n)
    e64 = rel(d.matvec(x, "fp64"), orc.matvec(x))
    e32 = rel(d.matvec(x, "fp32_resid"), orc.matvec(x))
    b = orc.rhs()
    xs, st = d.solve(b, tol=1e-10, mode="mixed")
    E = orc.energies(xs).sum()
    E_ref = orc.energies(np.linalg.solve(orc.assemble(), b)).sum()
    d.close()
    print("curved 2,562 patches: mat-vec rel err fp64", e64, "fp32_resid", e32, "energy", E, "LU", E_ref)
    assert e64 < 1e-12 and e32 < 3e-8
    assert abs(E - E_ref) / abs(E_ref) < 1e-7


def test_curved_points_generated_on_the_device():
    """The device generates curved (PNG1) point sets from the placed control
    points: a rotated and shifted curved instance beside a planar one gives
    the mat-vecs and energy coefficients of the host's point sets to
    rounding (the surface and its tangents are rigid-motion covariant)."""
    from beep3_amd import Mesh
    from beep3_amd.scene import Scene

    curved = Mesh.icosphere(4, 2.0).set_charges([[0.3, 0.0, -0.2, 1.0]]).set_curved(True)
    flat = Mesh.icosphere(3, 1.5).set_charges([[0.0, 0.1, 0.0, -1.0]])
    sc = Scene(80.0, 0.125, 6, 4, 1500)
    sc.add_instance(curved, 2.0, quat=[0.6, -0.2, 0.7, 0.3], offset=[1.0, -2.0, 0.5])
    sc.add_instance(flat, 2.0, offset=[6.0, 0.0, 0.0])
    sc.build()
    a, surf = sc.arrays(), sc.surface()
    assert surf["ctl"].shape[0] > 0 and np.any(surf["ctl_idx"] < 0)
    devs = []
    for mode in ("host", "device"):
        d = Device(0)
        d.set_far_field("direct")
        if mode == "host":
            d.set_geometry(a)
        else:
            d.set_surface(a, surf, 4, 6)
        d.set_tree(a)
        d.precompute(0.125, "fp64")
        d.precompute(0.125, "fp32")
        devs.append(d)
    x = np.random.default_rng(12).normal(size=2 * a["area"].shape[0])
    e64 = rel(devs[1].matvec(x, "fp64"), devs[0].matvec(x, "fp64"))
    e32 = rel(devs[1].matvec(x, "fp32"), devs[0].matvec(x, "fp32"))
    ef = [d.energy_coefficients(a["inst_patch_ptr"], a["inst_charge_ptr"], a["charges"]) for d in devs]
    ee = max(rel(ef[1][k], ef[0][k]) for k in range(2))
    print(f"curved device vs host points: fp64 {e64:.2e}, fp32 {e32:.2e}, energy coefficients {ee:.2e}")
    assert e64 < 1e-12 and e32 < 1e-6 and ee < 1e-12
    for d in devs:
        d.close()
