"""Pose batches (Python view of the beep_poses API in include/beep_serve.h).

    pb = PoseBatch(mesh, slots=8)            # the mesh's own near field, once
    res = pb.solve([(quat, offset), ...])    # instance 1 posed; instance 0 at the origin
    pb.close()

The mesh's own explicit pairs, their coefficients and the singular self
terms are computed once and shared by both instances of every pose (rigid
motions leave them unchanged); per pose only the pairs between the two
instances closer than the cutoff are integrated, everything else goes through
the far field.  Poses run concurrently on `slots` device contexts.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import BeepError, PoseOpts, PoseResult, dp, lib, vp

SOLVE_MODES = {"fp64": 0, "mixed": 1, "fp32": 2}
FAR_MODES = {"auto": -1, "direct": 0, "fmm": 1, "pw": 2}


class PoseBatch:
    def __init__(self, mesh, Dsolvent=80.0, kappa=0.125, quad=6, qual=0, nbsize=1500, Dprotein=2.0,
                 slots=8, mode="mixed", tol=1e-6, max_iter=100, max_refine=4, inner_tol=0.0,
                 far_field="fmm", fmm_order=5, fmm_leaf=64, pair_cutoff=0.0, gram_schmidt=1,
                 keep_fh=False, device=0):
        o = PoseOpts()
        lib().beep_pose_default_opts(C.byref(o))
        o.device, o.slots, o.solve_mode = int(device), int(slots), SOLVE_MODES[mode]
        o.tol, o.max_iter, o.max_refine, o.inner_tol = float(tol), int(max_iter), int(max_refine), float(inner_tol)
        o.far_mode, o.fmm_order, o.fmm_leaf = FAR_MODES[far_field], int(fmm_order), int(fmm_leaf)
        o.pair_cutoff, o.gram_schmidt, o.keep_fh = float(pair_cutoff), int(gram_schmidt), 1 if keep_fh else 0
        h = vp()
        if lib().beep_poses_create(C.byref(o), mesh._h, float(Dsolvent), float(kappa), int(quad), int(qual),
                                   int(nbsize), float(Dprotein), C.byref(h)) != 0:
            raise BeepError(lib().beep_host_last_error().decode())
        self._h = h
        self._mesh = mesh  # must outlive the batch
        self.n_patches = self.info()["n_patches"]

    def info(self) -> dict:
        n, pairs, cut, ms = C.c_int(), C.c_longlong(), C.c_double(), C.c_double()
        lib().beep_poses_info(self._h, C.byref(n), C.byref(pairs), C.byref(cut), C.byref(ms))
        return {"n_patches": n.value, "intra_pairs": pairs.value, "cutoff": cut.value, "setup_ms": ms.value}

    def solve(self, poses):
        """poses: sequence of (quaternion (w, x, y, z), centre offset); returns
        one dict per pose."""
        n = len(poses)
        q = np.ascontiguousarray([p[0] for p in poses], dtype=np.float64).reshape(n, 4)
        t = np.ascontiguousarray([p[1] for p in poses], dtype=np.float64).reshape(n, 3)
        res = (PoseResult * max(n, 1))()
        if lib().beep_poses_solve(self._h, n, q.ctypes.data_as(dp), t.ctypes.data_as(dp), res) != 0:
            raise BeepError("pose solve failed")
        out = []
        for k in range(n):
            d = {f: getattr(res[k], f) for f, _ in PoseResult._fields_}
            d["message"] = res[k].message.decode(errors="replace")
            out.append(d)
        return out

    def take_fh(self, k: int):
        f, h = np.empty(2 * self.n_patches), np.empty(2 * self.n_patches)
        if lib().beep_poses_take_fh(self._h, int(k), f.ctypes.data_as(dp), h.ctypes.data_as(dp)) != 0:
            raise BeepError(lib().beep_host_last_error().decode())
        return f, h

    def close(self):
        if getattr(self, "_h", None):
            lib().beep_poses_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
