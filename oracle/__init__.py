"""FP64 CPU oracle for the BEM mat-vec path -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker.  The product
package ``beep3_amd`` never imports it.

``liboracle.so`` (oracle/oracle.c) restates the reference functions listed in
its header; this module wraps it with ctypes and adds the closed forms used to
pin it:

* Born / Born-Kirkwood solvation energy of a central charge in a sphere
  (tools/born_ion.py:6-11 without kappa; with the salt factor 1/(1+kappa a)),
  energy unit conversion from src/bem/constants.h:14-51;
* Kirkwood's series for off-centre charges in a salted dielectric sphere
  (tools/off-centre.py:29-98, generalised to several charges).
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

# src/bem/constants.h:14-51
AVOGADRO = 6.02214179e23
EPSILON0 = 8.85418782e-12
ELEMENTARY_CHARGE = 1.60217646e-19
ANGSTROM = 1.0e-10
CONVERT_ENERGY_TO_KJ_PER_MOL = AVOGADRO * ELEMENTARY_CHARGE**2 / (EPSILON0 * ANGSTROM * 1000.0)


def build():
    subprocess.run(["make", "-s", "-B", "-C", _HERE], check=True)


def _load():
    if not os.path.exists(_LIB):
        build()
    return C.CDLL(_LIB)


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int)
        g = C.POINTER(OrcGeom)
        _lib.orc_pair.argtypes = [g, C.c_double, C.c_int, C.c_int, dp]
        _lib.orc_pair_1pt.argtypes = [g, C.c_double, C.c_int, C.c_int, dp]
        _lib.orc_singular.argtypes = [g, C.c_double, dp]
        _lib.orc_matvec_rows.argtypes = [g, C.c_double, dp, C.c_int, dp, C.c_int, ip, dp]
        _lib.orc_assemble_tree.argtypes = [g, C.c_double, dp, dp]
        _lib.orc_matvec_all.argtypes = [g, C.c_double, dp, dp, dp]
        _lib.orc_rhs.argtypes = [g, C.c_int, dp, C.c_double, dp]
        _lib.orc_energy_coeffs.argtypes = [g, C.c_int, C.c_int, dp, C.c_int, C.c_int, dp, dp]
        _lib.orc_gmres_dense.argtypes = [C.c_int, dp, dp, dp, C.c_double, C.c_int, dp]
        _lib.orc_gmres_dense.restype = C.c_int
        _lib.cref_create.argtypes = [g, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, dp]
        _lib.cref_create.restype = C.c_void_p
        _lib.cref_matvec.argtypes = [C.c_void_p, dp, dp]
        _lib.cref_matvec.restype = C.c_double
        _lib.cref_gmres.argtypes = [C.c_void_p, dp, dp, C.c_double, C.c_int, dp]
        _lib.cref_gmres.restype = C.c_int
        _lib.cref_destroy.argtypes = [C.c_void_p]
        pp = C.POINTER(ip)
        _lib.orc_octree.argtypes = [C.c_int, dp, dp, C.c_double, C.c_int, ip, C.POINTER(C.c_longlong),
                                    pp, pp, pp, pp]
        _lib.orc_free.argtypes = [C.c_void_p]
    return _lib


def octree_csr(centroid, cube_centre, cube_edge: float, nbsize: int) -> dict:
    """Leaves and neighbourhood contents of the adaptive octree over the
    patch centroids (octree_ref.cpp, restating src/fmm/octree.h under the
    policy of src/beep/beep.cpp:1585-1641), as CSR arrays."""
    pts = np.ascontiguousarray(centroid, dtype=np.float64).reshape(-1, 3)
    cc = np.ascontiguousarray(cube_centre, dtype=np.float64).reshape(3)
    nl, nn = C.c_int(), C.c_longlong()
    ptrs = [C.POINTER(C.c_int)() for _ in range(4)]
    rc = lib().orc_octree(pts.shape[0], _dp(pts), _dp(cc), float(cube_edge), int(nbsize), C.byref(nl),
                          C.byref(nn), *[C.byref(p) for p in ptrs])
    if rc != 0:
        raise RuntimeError("oracle octree failed")
    sizes = [nl.value + 1, pts.shape[0], nl.value + 1, nn.value]
    out = {}
    for name, p, k in zip(("leaf_ptr", "leaf_idx", "nb_ptr", "nb_idx"), ptrs, sizes):
        out[name] = np.ctypeslib.as_array(p, shape=(k,)).copy()
        lib().orc_free(C.cast(p, C.c_void_p))
    return out


class OrcGeom(C.Structure):
    _fields_ = [
        ("n", C.c_int),
        ("centroid", C.c_void_p), ("normal", C.c_void_p), ("area", C.c_void_p),
        ("eps", C.c_void_p), ("gc", C.c_void_p),
        ("qual_ptr", C.c_void_p), ("qual", C.c_void_p),
        ("quad_ptr", C.c_void_p), ("quad", C.c_void_p),
        ("gal_ptr", C.c_void_p), ("gal", C.c_void_p),
        ("n_leaves", C.c_int),
        ("leaf_ptr", C.c_void_p), ("leaf_idx", C.c_void_p),
        ("nb_ptr", C.c_void_p), ("nb_idx", C.c_void_p),
    ]


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


class Oracle:
    """Oracle bound to one scene's arrays (a dict as returned by
    ``beep3_amd.Scene.arrays()``)."""

    def __init__(self, arrays: dict, kappa: float, Dsolvent: float):
        self.a = {k: np.ascontiguousarray(v) for k, v in arrays.items()}
        self.kappa = float(kappa)
        self.Dsolvent = float(Dsolvent)
        a = self.a
        self.n = int(a["area"].shape[0])
        for kind in ("qual", "quad", "gal"):
            if a[kind].shape[0] != int(a[kind + "_ptr"][-1]):
                raise ValueError(f"the oracle needs the host point sets ({kind} is missing: "
                                 "build the scene with host points)")
        g = OrcGeom()
        g.n = self.n
        for name in ("centroid", "normal", "area", "eps", "gc", "qual_ptr", "qual", "quad_ptr",
                     "quad", "gal_ptr", "gal", "leaf_ptr", "leaf_idx", "nb_ptr", "nb_idx"):
            setattr(g, name, a[name].ctypes.data)
        g.n_leaves = int(a["leaf_ptr"].shape[0] - 1)
        self.g = g
        self._diag = None

    def pair(self, i, j, one_point=False):
        out = np.zeros(4)
        f = lib().orc_pair_1pt if one_point else lib().orc_pair
        f(C.byref(self.g), self.kappa, int(i), int(j), _dp(out))
        return out

    def diag(self):
        if self._diag is None:
            d = np.zeros((self.n, 4))
            lib().orc_singular(C.byref(self.g), self.kappa, _dp(d))
            self._diag = d
        return self._diag

    NEAR, FAR, DIAG = 1, 2, 4

    def matvec(self, x, rows=None, flags=7):
        """y = K x following the octree split; returns (f rows, h rows) stacked
        as a 2 x len(rows) array when rows is given, else the 2N vector."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        d = self.diag()
        if rows is None:
            y = np.zeros(2 * self.n)
            lib().orc_matvec_rows(C.byref(self.g), self.kappa, _dp(d), flags, _dp(x), self.n,
                                  None, _dp(y))
            return y
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        y = np.zeros(2 * len(rows))
        lib().orc_matvec_rows(C.byref(self.g), self.kappa, _dp(d), flags, _dp(x), len(rows),
                              _ip(rows), _dp(y))
        return y.reshape(2, len(rows))

    def matvec_all(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(2 * self.n)
        lib().orc_matvec_all(C.byref(self.g), self.kappa, _dp(self.diag()), _dp(x), _dp(y))
        return y

    def assemble(self):
        K = np.zeros((2 * self.n, 2 * self.n))
        lib().orc_assemble_tree(C.byref(self.g), self.kappa, _dp(self.diag()), _dp(K))
        return K

    def rhs(self):
        q = self.a["charges"]
        b = np.zeros(2 * self.n)
        lib().orc_rhs(C.byref(self.g), int(q.shape[0]), _dp(q), self.Dsolvent, _dp(b))
        return b

    def energy_coeffs(self):
        ecf = np.zeros(self.n)
        ech = np.zeros(self.n)
        pp, cp = self.a["inst_patch_ptr"], self.a["inst_charge_ptr"]
        for k in range(len(pp) - 1):
            lib().orc_energy_coeffs(C.byref(self.g), int(pp[k]), int(pp[k + 1]),
                                    _dp(self.a["charges"]), int(cp[k]), int(cp[k + 1]),
                                    _dp(ecf), _dp(ech))
        return ecf, ech

    def energies(self, x):
        """Per-instance solvation energy in kJ/mol (Mesh::calculate_energy,
        src/bem/mesh.cpp:1373-1403)."""
        ecf, ech = self.energy_coeffs()
        f, h = x[: self.n], x[self.n:]
        pp = self.a["inst_patch_ptr"]
        out = []
        for k in range(len(pp) - 1):
            s = slice(int(pp[k]), int(pp[k + 1]))
            e = np.sum(h[s] * self.a["eps"][s] * ech[s] - f[s] * ecf[s])
            out.append(0.5 * e * CONVERT_ENERGY_TO_KJ_PER_MOL)
        return np.array(out)

    @staticmethod
    def gmres_dense(K, b, tol=1e-6, max_iter=100, x0=None):
        n = b.shape[0]
        K = np.ascontiguousarray(K, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
        hist = np.zeros(max_iter)
        its = lib().orc_gmres_dense(n, _dp(K), _dp(b), _dp(x), tol, max_iter, _dp(hist))
        return x, its, hist[: its - 1]


class CpuReference:
    """The reference's CPU mat-vec path restated in C++/OpenMP
    (oracle/cpu_ref.cpp): FP32 24-byte LocalIntegrations built once and
    applied by scatter (beep.cpp:768-956, local_integrations.h:123-130), an
    FP64 plane-wave FMM of order 9 with the 9-node quadrature
    (fmm_bem_hybrid.h:184-642, the reference's <9,9,67>), and the reference
    GMRES (beep.cpp:1262-1483).  Timed as the CPU baseline."""

    def __init__(self, orc: "Oracle", order=9, nodes=9, leaf_target=40, stride=1):
        """stride > 1: coefficients of every stride-th near-list leaf only
        (a timing sample; the other records keep their indices, zero values)."""
        self.orc = orc
        self.n = orc.n
        info = np.zeros(9)
        self._h = lib().cref_create(C.byref(orc.g), orc.kappa, int(order), int(nodes), int(leaf_target),
                                    int(stride), _dp(info))
        self.info = {"setup_ms": info[0], "near_ms": info[1], "records": int(info[2]),
                     "residual_pairs": int(info[3]), "leaf_level": int(info[4]), "boxes": int(info[5]),
                     "tables_ms": info[6], "threads": int(info[7]), "records_computed": int(info[8])}

    def matvec(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty(2 * self.n)
        ms = lib().cref_matvec(self._h, _dp(x), _dp(y))
        return y, ms

    def gmres(self, b, tol=1e-6, max_iter=100, x0=None):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(2 * self.n) if x0 is None else np.array(x0, dtype=np.float64)
        ms = C.c_double()
        mv = lib().cref_gmres(self._h, _dp(b), _dp(x), float(tol), int(max_iter), C.byref(ms))
        return x, mv, ms.value

    def close(self):
        if self._h:
            lib().cref_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def born_energy(q, a, eps_in, eps_out, kappa=0.0):
    """Born(-Kirkwood) energy of a central point charge in a sphere, kJ/mol:
    -q^2/(8 pi a) (1/eps_in - 1/(eps_out (1 + kappa a))) * conversion."""
    return -(q * q) / (8.0 * math.pi * a) * (
        1.0 / eps_in - 1.0 / (eps_out * (1.0 + kappa * a))) * CONVERT_ENERGY_TO_KJ_PER_MOL


def kirkwood_coefficient(n, a, eps_in, eps_out, kappa):
    """Reaction-potential coefficient c_n of a dielectric sphere (radius a,
    eps_in) in a salted solvent (eps_out, kappa): the reaction potential at r
    of a unit charge at r0 is sum_n c_n r^n r0^n P_n(cos g) (potential q/(eps r)
    convention).  From continuity of phi and eps dphi/dr at r = a with the
    exterior solution k_n(kappa r) (modified spherical Bessel, 2nd kind):
        c_n = a^-(2n+1) [(n+1) + eps_out g/eps_in] / (n eps_in - eps_out g),
        g = x k_n'(x)/k_n(x), x = kappa a;  g -> -(n+1) as kappa -> 0."""
    if kappa == 0.0:
        g = -(n + 1.0)
    else:
        from scipy.special import spherical_kn
        x = kappa * a
        g = x * spherical_kn(n, x, derivative=True) / spherical_kn(n, x)
    return a ** -(2 * n + 1) * ((n + 1) + eps_out * g / eps_in) / (n * eps_in - eps_out * g)


def kirkwood_energy(xyzq, a, eps_in, eps_out, kappa, nmax=80):
    """Reaction-field (solvation) energy in kJ/mol of point charges inside a
    sphere centred at the origin: Kirkwood's series (J. Chem. Phys. 2, 351
    (1934)), as tools/off-centre.py:29-98 evaluates for one charge, summed over
    all charge pairs: W = 1/2 conv/(4 pi) sum_n c_n sum_ij q_i q_j (r_i r_j)^n
    P_n(cos g_ij).  n = 0 alone is the Born-Kirkwood energy."""
    from scipy.special import eval_legendre

    xyzq = np.asarray(xyzq, dtype=np.float64)
    pos, q = xyzq[:, :3], xyzq[:, 3]
    r = np.linalg.norm(pos, axis=1)
    rr = np.outer(r, r)
    with np.errstate(invalid="ignore", divide="ignore"):
        cosg = np.where(rr > 0, (pos @ pos.T) / np.where(rr > 0, rr, 1.0), 1.0)
    cosg = np.clip(cosg, -1.0, 1.0)
    qq = np.outer(q, q)
    total = 0.0
    for n in range(nmax + 1):
        cn = kirkwood_coefficient(n, a, eps_in, eps_out, kappa)
        total += cn * np.sum(qq * rr**n * eval_legendre(n, cosg))
    return 0.5 * total * CONVERT_ENERGY_TO_KJ_PER_MOL / (4.0 * math.pi)
