"""The M2L symmetry of the Chebyshev FMM (host side, no GPU): for every
well-separated offset t in [-3, 3]^3 the library's canonical offset t_c and
node map psi satisfy K_t[psi(m)][psi(n)] = K_{t_c}[m][n], with the kernel
tables computed here from their definition G(|t + (xi_m - xi_n)/2| h) for the
Laplace and the Yukawa kernel.  A wrong axis, sign or node index in the map
breaks the equality for some offset."""
import ctypes as C
import itertools

import numpy as np
import pytest

from beep3_amd._lib import lib


def tables(t, P, kappa_h):
    xi = np.cos((2 * np.arange(P) + 1) * np.pi / (2 * P))
    idx = np.array(list(itertools.product(range(P), repeat=3)))  # node m = (a P + b) P + c
    X = xi[idx]
    v = np.asarray(t, float)[None, None, :] + 0.5 * (X[:, None, :] - X[None, :, :])
    r = np.linalg.norm(v, axis=2)
    return 1.0 / r, np.exp(-kappa_h * r) / r


def symmetry(t, P):
    c = (C.c_int * 3)()
    psi = (C.c_int * (P ** 3))()
    k = lib().beep_hip_fmm_symmetry(t[0], t[1], t[2], P, c, psi)
    assert k >= 0
    return k, tuple(c), np.array(psi[:], dtype=np.int64)


@pytest.mark.parametrize("P", [3, 4, 5, 6])
def test_canonical_tables_reproduce_every_offset(P):
    seen = {}
    worst = 0.0
    for t in itertools.product(range(-3, 4), repeat=3):
        if max(abs(x) for x in t) < 2:
            continue
        k, tc, psi = symmetry(t, P)
        assert tc[0] >= tc[1] >= tc[2] >= 0 and tc[0] in (2, 3)
        assert sorted(abs(x) for x in t) == sorted(tc)
        assert sorted(psi) == list(range(P ** 3))  # a permutation of the nodes
        seen.setdefault(k, tc)
        assert seen[k] == tc
        for Kt, Kc in zip(tables(t, P, 0.7), tables(tc, P, 0.7)):
            worst = max(worst, float(np.abs(Kt[np.ix_(psi, psi)] - Kc).max() / np.abs(Kc).max()))
    assert len(seen) == 16
    assert worst < 1e-14


def test_near_offsets_are_rejected():
    c = (C.c_int * 3)()
    psi = (C.c_int * 125)()
    assert lib().beep_hip_fmm_symmetry(1, 0, -1, 5, c, psi) < 0
    assert lib().beep_hip_fmm_symmetry(4, 0, 0, 5, c, psi) < 0
