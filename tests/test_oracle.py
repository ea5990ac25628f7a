"""Pins of the FP64 CPU oracle (oracle/oracle.c) against closed forms,
invariants and independent numerics.  CPU only."""
import math

import numpy as np
import pytest

import oracle
from beep3_amd import Mesh, Scene


def make(sub=3, radius=2.0, kappa=0.125, quad=6, qual=0, Dp=2.0, charges=((0, 0, 0, 1.0),),
         nbsize=1500):
    m = Mesh.icosphere(sub, radius).set_charges(np.array(charges, float))
    sc = Scene(80.0, kappa, quad, qual, nbsize).add_instance(m, Dp).build()
    return sc, oracle.Oracle(sc.arrays(), kappa, 80.0)


def solve_energy(o):
    K = o.assemble()
    x = np.linalg.solve(K, o.rhs())
    return o.energies(x).sum(), K, x


def test_born_kirkwood_converges():
    """Central +1e in a sphere (eps 2 | 80, kappa 0.125): the BEM energy
    approaches the Born-Kirkwood closed form as the mesh is refined."""
    born = oracle.born_energy(1.0, 2.0, 2.0, 80.0, 0.125)
    assert born == pytest.approx(-170.196, abs=1e-3)
    errs = []
    for sub in (2, 3, 4):
        _, o = make(sub)
        E, _, _ = solve_energy(o)
        errs.append(abs(E - born) / abs(born))
    assert errs[2] < 5e-3
    assert errs[0] > errs[1] > errs[2]


def test_born_without_salt():
    born = oracle.born_energy(1.0, 2.0, 2.0, 80.0, 0.0)
    assert born == pytest.approx(-169.328, abs=1e-3)
    _, o = make(3, kappa=0.0)
    E, _, _ = solve_energy(o)
    assert abs(E - born) / abs(born) < 1.5e-2


def test_double_layer_identity():
    """kappa = 0, f = 1, h = 0: the double-layer integral over a closed
    surface is -1/2 at a surface point, so B row sums plus the geometric
    correction (1/2)(1 + 1/eps) give y_f = 1; D vanishes (kappa guard)."""
    devs = []
    for sub in (2, 3):
        _, o = make(sub, kappa=0.0)
        x = np.concatenate([np.ones(o.n), np.zeros(o.n)])
        y = o.matvec(x)
        devs.append(np.max(np.abs(y[: o.n] - 1.0)))
        assert np.max(np.abs(y[o.n:])) < 1e-12
    assert devs[1] < 2e-2
    assert devs[1] < devs[0]


def _two_points(kappa, eps_j=40.0):
    """A two-patch geometry with single points, to probe the integrand."""
    rng = np.random.default_rng(7)
    c = np.array([[0.0, 0.0, 0.0], [1.3, -0.4, 0.7]])
    n = rng.normal(size=(2, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    a = {"centroid": c, "normal": n, "area": np.array([0.3, 0.2]), "eps": np.array([40.0, eps_j]),
         "gc": np.full(2, 0.5)}
    pts = np.hstack([c, n, np.array([[1.0], [0.2]])])  # qual weight 1 / quad weight = area
    for k in ("qual", "quad", "gal"):
        a[k + "_ptr"] = np.array([0, 1, 2], dtype=np.int32)
        a[k] = pts.copy()
    a["qual"][:, 6] = 1.0
    a["leaf_ptr"] = np.array([0, 2], dtype=np.int32)
    a["leaf_idx"] = np.array([0, 1], dtype=np.int32)
    a["nb_ptr"] = np.array([0, 2], dtype=np.int32)
    a["nb_idx"] = np.array([0, 1], dtype=np.int32)
    return a, oracle.Oracle(a, kappa, 80.0)


def _G(x, y, kappa):
    r = np.linalg.norm(y - x)
    return 1.0 / (4 * math.pi * r), math.exp(-kappa * r) / (4 * math.pi * r)


def test_integrand_against_finite_differences():
    """A, B, C, D of one point pair against central differences of the bare
    kernels G0 = 1/(4 pi r) and u = e^{-kr}/(4 pi r)."""
    kappa, eps_j = 0.37, 25.0
    a, o = _two_points(kappa, eps_j)
    A, B, C, D = o.pair(0, 1)
    x, y = a["centroid"][0], a["centroid"][1]
    n0, n = a["normal"][0], a["normal"][1]
    w = 0.2
    h = 1e-5

    def dn_y(fun, x, y, n):
        return (fun(x, y + h * n) - fun(x, y - h * n)) / (2 * h)

    g0 = lambda x, y: _G(x, y, kappa)[0]
    u = lambda x, y: _G(x, y, kappa)[1]
    assert A == pytest.approx(w * (g0(x, y) - u(x, y)), rel=1e-12)
    B_fd = w * (dn_y(g0, x, y, n) / eps_j - dn_y(u, x, y, n))
    assert B == pytest.approx(B_fd, rel=1e-7)
    # C = -n0 . (grad_y G0 - grad_y u / eps) = n0 . grad_x (G0 - u/eps)
    dn_x = lambda fun: (fun(x + h * n0, y) - fun(x - h * n0, y)) / (2 * h)
    C_fd = w * (dn_x(g0) - dn_x(u) / eps_j)
    assert C == pytest.approx(C_fd, rel=1e-7)
    # D = -(1/eps) n0 . grad_y grad_y (G0 - u) . n  (second derivative in y)
    hh = 1e-4
    diff = lambda xx, yy: g0(xx, yy) - u(xx, yy)
    d2 = (diff(x, y + hh * n + hh * n0) - diff(x, y + hh * n - hh * n0)
          - diff(x, y - hh * n + hh * n0) + diff(x, y - hh * n - hh * n0)) / (4 * hh * hh)
    D_fd = -w * d2 / eps_j
    assert D == pytest.approx(D_fd, rel=1e-5)


def test_kappa_zero_limits():
    a, o = _two_points(0.0, 25.0)
    A, B, C, D = o.pair(0, 1)
    assert A == 0.0 and D == 0.0
    a2, o2 = _two_points(1e-7, 25.0)
    A2, B2, C2, D2 = o2.pair(0, 1)
    assert B2 == pytest.approx(B, rel=1e-6) and C2 == pytest.approx(C, rel=1e-6)
    assert abs(A2) < 1e-7


def test_gmres_matches_dense_solve():
    """GMRES with the reference conventions (beep.cpp:1262-1483) converges
    to the LU solution of the same dense operator."""
    _, o = make(3)
    K = o.assemble()
    b = o.rhs()
    x_lu = np.linalg.solve(K, b)
    x, its, hist = oracle.Oracle.gmres_dense(K, b, 1e-10, 100)
    assert np.linalg.norm(x - x_lu) / np.linalg.norm(x_lu) < 1e-8
    assert its <= 30
    assert np.all(np.diff(hist) <= 1e-15)  # residual estimates never increase


def test_rhs_closed_form():
    """Single charge at the centre with centroid collocation: phi = q/|c| and
    grad phi . n = -q (n . c)/|c|^3, scaled by 1/(4 pi Dsolvent)."""
    sc, o = make(2)
    a = sc.arrays()
    b = o.rhs()
    c, n = a["centroid"], a["normal"]
    r = np.linalg.norm(c, axis=1)
    f_ref = 1.0 / r / (4 * math.pi * 80.0)
    h_ref = -np.sum(n * c, axis=1) / r**3 / (4 * math.pi * 80.0)
    assert np.allclose(b[: o.n], f_ref, rtol=1e-13)
    assert np.allclose(b[o.n:], h_ref, rtol=1e-12)


def test_energy_coefficients_gauss_law():
    sc, o = make(3, charges=((0.3, -0.2, 0.1, 0.7), (-0.4, 0.1, 0.5, -0.2)))
    ecf, ech = o.energy_coeffs()
    net = 0.7 - 0.2
    assert np.sum(ecf) == pytest.approx(-net, rel=1e-12)


def test_kirkwood_series_reduces_to_born():
    q = np.array([[0.0, 0.0, 0.0, 1.0]])
    for kappa in (0.0, 0.125):
        assert oracle.kirkwood_energy(q, 2.0, 2.0, 80.0, kappa) == pytest.approx(
            oracle.born_energy(1.0, 2.0, 2.0, 80.0, kappa), rel=1e-12)


def test_off_centre_charges_match_kirkwood():
    """Off-centre charges in a salted dielectric sphere (the config-2 physics
    on a coarse mesh): BEM energy vs Kirkwood's series."""
    rng = np.random.default_rng(2)
    pts = []
    while len(pts) < 5:
        p = rng.uniform(-1, 1, 3)
        if np.linalg.norm(p) < 1:
            pts.append(list(p * 12.0) + [rng.choice([-1, 1]) * rng.uniform(0.25, 1.0)])
    q = np.array(pts)
    m = Mesh.icosphere(4, 20.0).set_charges(q)
    sc = Scene(80.0, 0.1, 6, 0, 1500).add_instance(m, 4.0).build()
    o = oracle.Oracle(sc.arrays(), 0.1, 80.0)
    E, _, _ = solve_energy(o)
    ref = oracle.kirkwood_energy(q, 20.0, 4.0, 80.0, 0.1)
    assert abs(E - ref) / abs(ref) < 2e-2


def test_tree_split_vs_all_pairs():
    """D_tree (explicit near pairs, 1-point far pairs) against D_all (full
    rules for every pair): they differ only in far-pair quadrature."""
    sc, o = make(3, nbsize=60)
    x = np.random.default_rng(1).normal(size=2 * o.n)
    y_tree = o.matvec(x)
    y_all = o.matvec_all(x)
    near_only = o.matvec(x, flags=oracle.Oracle.NEAR | oracle.Oracle.DIAG)
    rel = np.linalg.norm(y_tree - y_all) / np.linalg.norm(y_all)
    assert rel < 5e-3
    # the far part is not negligible, so the split is exercised
    assert np.linalg.norm(y_tree - near_only) / np.linalg.norm(y_tree) > 1e-3
