"""Solve server, CPU side: option handling, ownership of submitted scenes and
the failure path (on a machine without a GPU every request comes back as a
failed result with a message; nothing hangs or crashes)."""
import os
import time

import pytest

from beep3_amd._lib import BeepError
from beep3_amd.scene import Mesh, Scene
from beep3_amd.serve import Server


def _has_gpu():
    import ctypes as C

    from beep3_amd._lib import lib, vp

    h = vp()
    if lib().beep_hip_create(0, C.byref(h)) == 0:
        lib().beep_hip_destroy(h)
        return True
    return False


def test_bad_options_are_rejected():
    with pytest.raises(BeepError):
        Server(slots=0)
    with pytest.raises(BeepError):
        Server(budget_gib=-1)


def test_submitted_scene_changes_owner():
    srv = Server(slots=1)
    m = Mesh.icosphere(2, 2.0).set_charges([[0, 0, 0, 1.0]])
    sc = Scene(80.0, 0.125, 6, 0, 1500).add_instance(m, 2.0)
    srv.submit(sc)
    with pytest.raises(ValueError):
        srv.submit(sc)
    srv.close()


@pytest.mark.skipif(_has_gpu(), reason="failure path needs a machine without a GPU")
def test_requests_fail_cleanly_without_gpu():
    srv = Server(slots=3)
    meshes = [Mesh.icosphere(2, 2.0).set_charges([[0, 0, 0, 1.0]]) for _ in range(5)]
    ids = [srv.submit(Scene(80.0, 0.125, 6, 0, 1500).add_instance(m, 2.0)) for m in meshes]
    got = {}
    t0 = time.time()
    while len(got) < len(ids) and time.time() - t0 < 60:
        r = srv.poll(timeout=1.0)
        if r is not None:
            got[r["id"]] = r
    assert sorted(got) == sorted(ids)
    for r in got.values():
        assert r["status"] < 0
        assert "device context" in r["message"]
        assert r["n_patches"] == 162
        assert r["t_done"] >= r["t_submit"]
    assert srv.pending() == 0
    srv.close()


def test_oversized_request_is_refused():
    """A request whose device estimate exceeds the whole budget fails
    instead of waiting forever (checked before the device is read)."""
    if _has_gpu():
        pytest.skip("CPU-only check")
    srv = Server(slots=1, budget_gib=1e-6)
    m = Mesh.icosphere(3, 2.0).set_charges([[0, 0, 0, 1.0]])
    rid = srv.submit(Scene(80.0, 0.125, 6, 0, 1500).add_instance(m, 2.0))
    r = srv.poll(timeout=30.0)
    assert r is not None and r["id"] == rid and r["status"] < 0
    assert "budget" in r["message"]
    srv.close()


def test_estimate_counts_stored_coefficients():
    """The admission estimate with stored coefficients exceeds the on-the-fly
    one by the coefficient bytes (16 per explicit pair in mixed mode)."""
    import numpy as np

    srv = Server(slots=1)
    m = Mesh.icosphere(3, 2.0).set_charges([[0, 0, 0, 1.0]])
    sc = Scene(80.0, 0.125, 6, 0, 1500).add_instance(m, 2.0)
    stored, fly = srv.estimate(sc)
    sc.build()
    a = sc.arrays()
    nnz = int(np.sum(np.diff(a["leaf_ptr"]) * np.diff(a["nb_ptr"])))
    assert fly > 0 and stored - fly == 16.0 * nnz
    srv.close()
