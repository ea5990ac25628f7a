"""Solve server with continuous batching (Python view of include/beep_serve.h).

    srv = Server(slots=8)               # budget_gib=0: the device's free memory
    rid = srv.submit(scene)          # a Scene with its instances, not built
    r = srv.poll(timeout=1.0)        # next finished request (dict) or None
    srv.close()

Each slot is a worker thread with a resident device context; requests are
admitted as slots and device memory free up and run concurrently on the GPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import BeepError, ServerOpts, ServerResult, lib, vp
from .scene import Scene

SOLVE_MODES = {"fp64": 0, "mixed": 1, "fp32": 2}
FAR_MODES = {"auto": -1, "direct": 0, "fmm": 1, "pw": 2}
NEAR_MODES = {"auto": -1, "precompute": 0, "on_the_fly": 1}


class Server:
    def __init__(self, slots=4, budget_gib=0.0, mode="mixed", tol=1e-6, max_iter=100,
                 max_refine=4, inner_tol=0.0, far_field="auto", fmm_order=5, fmm_leaf=64,
                 keep_fh=False, device=0, near_mode="auto", device_tree="auto"):
        o = ServerOpts()
        lib().beep_server_default_opts(C.byref(o))
        o.device, o.slots, o.hbm_budget = int(device), int(slots), float(budget_gib) * 2.0**30
        o.solve_mode = SOLVE_MODES[mode]
        o.tol, o.max_iter, o.max_refine, o.inner_tol = float(tol), int(max_iter), int(max_refine), float(inner_tol)
        o.far_mode, o.fmm_order, o.fmm_leaf = FAR_MODES[far_field], int(fmm_order), int(fmm_leaf)
        o.keep_fh = 1 if keep_fh else 0
        o.near_mode = NEAR_MODES[near_mode]
        o.device_tree = -1 if device_tree == "auto" else (1 if device_tree else 0)
        self._opts = o
        h = vp()
        if lib().beep_server_create(C.byref(o), C.byref(h)) != 0:
            raise BeepError(lib().beep_host_last_error().decode())
        self._h = h
        self._keep = {}  # request id -> meshes that must outlive the request

    def submit(self, scene: Scene) -> int:
        """Hand the (unbuilt) scene to the server; the Scene object gives up
        its handle."""
        if scene._h is None:
            raise ValueError("scene already submitted")
        rid = C.c_longlong()
        if lib().beep_server_submit(self._h, scene._h, C.byref(rid)) != 0:
            raise BeepError(lib().beep_server_last_error(self._h).decode())
        self._keep[rid.value] = list(scene._meshes)
        scene._h = None
        return rid.value

    def poll(self, timeout=0.0):
        r = ServerResult()
        rc = lib().beep_server_poll(self._h, float(timeout), C.byref(r))
        if rc == 1:
            return None
        if rc != 0:
            raise BeepError(lib().beep_server_last_error(self._h).decode())
        self._keep.pop(r.id, None)
        out = {f: getattr(r, f) for f, _ in ServerResult._fields_}
        out["message"] = r.message.decode(errors="replace")
        out["latency_s"] = r.t_done - r.t_submit
        return out

    def take_fh(self, rid: int, n: int):
        f, h = np.empty(n), np.empty(n)
        dp = C.POINTER(C.c_double)
        if lib().beep_server_take_fh(self._h, int(rid), f.ctypes.data_as(dp), h.ctypes.data_as(dp)) != 0:
            raise BeepError(lib().beep_server_last_error(self._h).decode())
        return f, h

    def estimate(self, scene: Scene):
        """(bytes with stored coefficients, bytes on the fly) that admission
        charges this scene (built here if it is not; it stays the caller's)."""
        st, fl = C.c_double(), C.c_double()
        if lib().beep_server_estimate(C.byref(self._opts), scene._h, C.byref(st), C.byref(fl)) != 0:
            raise BeepError(lib().beep_host_last_error().decode())
        return st.value, fl.value

    def clock(self) -> float:
        return lib().beep_server_clock(self._h)

    def pending(self) -> int:
        return lib().beep_server_pending(self._h)

    def close(self):
        if getattr(self, "_h", None):
            lib().beep_server_destroy(self._h)
            self._h = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
