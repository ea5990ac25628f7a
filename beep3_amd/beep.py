"""BEEP: the solver object of the reference (src/beep/beep.h:118-342), for the
hot path only, on one MI355X.

Mirrors the reference's Python usage (tools/run_pair.py:34-51):

    beep = BEEP(Dsolvent, kappa, quad_points, qual_points, nbsize)
    lib_id = beep.load_library_mesh(mesh)
    beep.insert_mesh_instance(lib_id, offset, rotation, Dprotein)
    beep.solve(tol, max_iter)
    energies = beep.calculate_energies()

``do_bem_fmm_iteration`` is the drop-in mat-vec (beep.cpp:1485-1583) and runs
through ``beep_hip_matvec``.  Every numerical step runs on the GPU; the
library raises if its HIP extension is missing.
"""
from __future__ import annotations

import time

import numpy as np

from .device import Device
from .scene import Mesh, Scene

# src/bem/constants.h:14-51
AVOGADRO = 6.02214179e23
EPSILON0 = 8.85418782e-12
ELEMENTARY_CHARGE = 1.60217646e-19
ANGSTROM = 1.0e-10
CONVERT_ENERGY_TO_KJ_PER_MOL = AVOGADRO * ELEMENTARY_CHARGE**2 / (EPSILON0 * ANGSTROM * 1000.0)


class BEEP:
    def __init__(self, Dsolvent=80.0, kappa=0.0, quad_points=0, qual_points=4, nbsize=1500,
                 device=0, precision="fp32", far_field="auto", fmm_order=5, fmm_leaf=48,
                 device_points=True, device_tree=False):
        self.Dsolvent, self.kappa = float(Dsolvent), float(kappa)
        self.quad, self.qual, self.nbsize = int(quad_points), int(qual_points), int(nbsize)
        self.precision = precision  # coefficient set precomputed at init
        self.device_id = device
        self.far_field, self.fmm_order, self.fmm_leaf = far_field, int(fmm_order), int(fmm_leaf)
        # point sets made on the device from the surface (else on the host and uploaded)
        self.device_points = bool(device_points)
        # octree and near list built on the device (True, or "auto": from 65,536
        # patches); the scene's arrays then hold no leaf lists (Device.tree_arrays)
        self.device_tree = device_tree
        self.library: list[Mesh] = []
        self.instances: list[tuple] = []
        self.scene: Scene | None = None
        self.dev: Device | None = None
        self.f = self.h = None
        self.run_info: dict = {}
        self._ready = False

    # ---- scene set-up (beep.cpp:197-290) ----
    def load_library_mesh(self, mesh: Mesh) -> int:
        self.library.append(mesh)
        return len(self.library) - 1

    def insert_mesh_instance(self, lib_id: int, offset=None, rotation=None, Dprotein=2.0) -> int:
        self.instances.append((lib_id, None if offset is None else np.asarray(offset, float),
                               None if rotation is None else np.asarray(rotation, float),
                               float(Dprotein)))
        self._ready = False
        return len(self.instances) - 1

    def clear_mesh_instances(self):
        self.instances.clear()
        self._ready = False

    def init(self, precisions=None):
        """Build patches and octree (the octree on the device for large
        scenes), upload to the GPU and precompute the explicit-pair
        coefficients (precalc_neighbour_interactions)."""
        t0 = time.perf_counter()
        dev_points = self.device_points
        n_items = sum(self.library[lib_id].info()["n_vertices"] for lib_id, _, _, _ in self.instances)
        dev_tree = n_items >= 65536 if self.device_tree == "auto" else bool(self.device_tree)
        sc = Scene(self.Dsolvent, self.kappa, self.quad, self.qual, self.nbsize,
                   host_points=not dev_points, tree=not dev_tree)
        for lib_id, off, rot, Dp in self.instances:
            sc.add_instance(self.library[lib_id], Dp, quat=rot, offset=off)
        sc.build()
        t1 = time.perf_counter()
        a = sc.arrays()
        if self.dev is None:
            self.dev = Device(self.device_id)
        self.dev.set_far_field(self.far_field, self.fmm_order, self.fmm_leaf)
        if dev_points:
            self.dev.set_surface(a, sc.surface(), self.qual, self.quad)
        else:
            self.dev.set_geometry(a)
        sz = sc.sizes()
        if dev_tree:
            n_leaves, _ = self.dev.build_tree(list(sz.cube_centre), sz.cube_edge, self.nbsize)
            t = self.dev.tree_arrays(lists=False)
            explicit = int(np.sum(np.diff(t["leaf_ptr"]).astype(np.int64) * np.diff(t["nb_ptr"])))
        else:
            self.dev.set_tree(a)
            n_leaves, explicit = sz.n_leaves, int(sz.n_near_pairs)
        t2 = time.perf_counter()
        for p in (precisions or [self.precision]):
            self.dev.precompute(self.kappa, p)
        self.dev.synchronize()
        t3 = time.perf_counter()
        self.scene = sc
        self.n = int(a["area"].shape[0])
        self.run_info.update(patches=self.n, nbsize=self.nbsize, leaves=n_leaves,
                             explicit=explicit, tree_depth=sz.tree_depth if not dev_tree else None,
                             host_setup_ms=(t1 - t0) * 1e3, upload_ms=(t2 - t1) * 1e3,
                             precalc_ms=(t3 - t2) * 1e3)
        self._ready = True

    def host_scene(self) -> Scene:
        """The same scene with its point sets built on the host (for CPU
        checks of the operator; the device path generates its own)."""
        sc = Scene(self.Dsolvent, self.kappa, self.quad, self.qual, self.nbsize, host_points=True)
        for lib_id, off, rot, Dp in self.instances:
            sc.add_instance(self.library[lib_id], Dp, quat=rot, offset=off)
        return sc.build()

    def _need(self):
        if not self._ready:
            self.init()

    # ---- hot path ----
    def do_bem_fmm_iteration(self, lhs, precision="fp64"):
        """result = K lhs, lhs = [f; h] (beep.cpp:1485-1583)."""
        self._need()
        return self.dev.matvec(lhs, precision)

    def set_rhs(self):
        self._need()
        a = self.scene.arrays()
        return self.dev.rhs(a["charges"], self.Dsolvent)

    def solve(self, tol=1e-6, max_iter=100, mode="fp64", **kw):
        """RHS + GMRES (beep.cpp:417-512).  mode: 'fp64' (plain), 'mixed'
        (FP32 mat-vecs inside FP64 refinement) or 'fp32'."""
        self._need()
        t0 = time.perf_counter()
        b = self.set_rhs()
        t1 = time.perf_counter()
        x0 = None
        if self.f is not None and kw.pop("warm_start", False):
            x0 = np.concatenate([self.f, self.h])
        x, st = self.dev.solve(b, x0, tol=tol, max_iter=max_iter, mode=mode, **kw)
        t2 = time.perf_counter()
        self.f, self.h = x[: self.n].copy(), x[self.n:].copy()
        self.run_info.update(rhs_ms=(t1 - t0) * 1e3, gmres_ms=(t2 - t1) * 1e3, **st)
        return self.run_info

    def reset_library_fh_vals(self):
        self.f = self.h = None

    # ---- energies (Mesh::calculate_energy, mesh.cpp:1373-1403) ----
    def energy_coefficients(self):
        a = self.scene.arrays()
        ef, eh = self.dev.energy_coefficients(a["inst_patch_ptr"], a["inst_charge_ptr"],
                                              a["charges"])
        # Gauss-law renormalisation (mesh.cpp:891-925), per instance
        pp, cp = a["inst_patch_ptr"], a["inst_charge_ptr"]
        for k in range(len(pp) - 1):
            s = slice(int(pp[k]), int(pp[k + 1]))
            net = float(np.sum(a["charges"][int(cp[k]):int(cp[k + 1]), 3]))
            total_f = -float(np.sum(ef[s]))
            err = net - total_f
            ef[s] -= (a["area"][s] / np.sum(a["area"][s])) * err
        return ef, eh

    def calculate_energies(self, per_instance=False):
        self._need()
        if self.f is None:
            raise RuntimeError("solve() first")
        a = self.scene.arrays()
        ef, eh = self.energy_coefficients()
        pp = a["inst_patch_ptr"]
        out = []
        for k in range(len(pp) - 1):
            s = slice(int(pp[k]), int(pp[k + 1]))
            e = np.sum(self.h[s] * a["eps"][s] * eh[s] - self.f[s] * ef[s])
            out.append(0.5 * e * CONVERT_ENERGY_TO_KJ_PER_MOL)
        out = np.array(out)
        return out if per_instance else float(out.sum())

    def write_fh(self, path):
        """'f h' per line with 12 significant digits (beep.cpp:539-570)."""
        with open(path, "w") as fh:
            for f, h in zip(self.f, self.h):
                fh.write(f"{f:.12g} {h:.12g}\n")

    def close(self):
        if self.dev is not None:
            self.dev.close()
            self.dev = None
        self._ready = False
