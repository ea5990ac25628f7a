"""CPU: the plane-wave FMM tables (beep3_amd/csrc/pw_tables.cpp) against
direct sums of the screened and plain Coulomb kernels, through the C ABI
diagnostic beep_hip_pw_selftest.  Every check compares with closed forms
(e^{-beta r}/r and its gradient for seeded charges and dipoles), so a wrong
sign, index, parity or normalisation in any table shows up as an O(1) error."""
import ctypes as C

import numpy as np
import pytest

from beep3_amd._lib import lib


def selftest(order, nodes, beta):
    e = np.zeros(5)
    rc = lib().beep_hip_pw_selftest(order, nodes, beta, e.ctypes.data_as(C.POINTER(C.c_double)))
    assert rc == 0
    return e


@pytest.mark.parametrize("beta", [0.0, 0.125, 0.5])
def test_order12_18nodes(beta):
    """Six-digit class (order 12, 18 quadrature nodes): measured 3e-6
    (quadrature), 1e-5..1.2e-4 (P2M..L2P over the six directions, worst
    target corner), 1e-6..2e-5 through M2M / L2L."""
    q, e1, e2, eg, rot = selftest(12, 18, beta)
    print(f"beta {beta}: quad {q:.2e} m2x/x2l {e1:.2e} m2m/l2l {e2:.2e} grad {eg:.2e} rot {rot:.1e}")
    assert q < 1e-5
    assert e1 < 3e-4
    assert e2 < 6e-5
    assert eg < 6e-3
    assert rot < 1e-12


def test_order9_9nodes():
    """Three-digit class (the reference's <9,9,67>): measured 4e-3
    (quadrature), 1.7e-2 worst corner, 7e-4 through M2M / L2L."""
    q, e1, e2, eg, rot = selftest(9, 9, 0.125)
    print(f"order 9: quad {q:.2e} m2x/x2l {e1:.2e} m2m/l2l {e2:.2e} grad {eg:.2e}")
    assert q < 1e-2 and e1 < 5e-2 and e2 < 3e-3 and rot < 1e-12


def test_accuracy_improves_with_order():
    lo = selftest(9, 9, 0.125)
    hi = selftest(12, 18, 0.125)
    assert hi[0] < lo[0] / 100 and hi[1] < lo[1] / 100 and hi[2] < lo[2] / 100


def test_bad_arguments_are_refused():
    e = np.zeros(5)
    dp = C.POINTER(C.c_double)
    assert lib().beep_hip_pw_selftest(12, 7, 0.1, e.ctypes.data_as(dp)) != 0
    assert lib().beep_hip_pw_selftest(40, 18, 0.1, e.ctypes.data_as(dp)) != 0
    assert lib().beep_hip_pw_selftest(12, 18, -1.0, e.ctypes.data_as(dp)) != 0
