"""Host-side meshes and scenes (Python view of include/beep_host.h).

``Mesh`` is a reference mesh with charges (the reference's Mesh, built here by
seeded generators instead of read from a .mtz), ``Scene`` is the set of rigid
instances with the solver parameters (the reference's BEEP constructor
arguments, src/beep/beep.h:118-160) plus the octree that fixes the explicit
pair list.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import SceneSizes, check_host, dp, ip, lib, vp


def _dp(a):
    return a.ctypes.data_as(dp)


def _ip(a):
    return a.ctypes.data_as(ip)


class Mesh:
    def __init__(self, handle):
        self._h = vp(handle) if not isinstance(handle, vp) else handle

    @classmethod
    def icosphere(cls, subdivisions: int, radius: float) -> "Mesh":
        h = vp()
        check_host(lib().beep_mesh_icosphere(subdivisions, radius, C.byref(h)))
        return cls(h)

    @classmethod
    def star(cls, subdivisions: int, r0: float, lmax: int = 8, sigma0: float = 0.04,
             seed: int = 0) -> "Mesh":
        h = vp()
        check_host(lib().beep_mesh_star(subdivisions, r0, lmax, sigma0, seed, C.byref(h)))
        return cls(h)

    @classmethod
    def read(cls, path: str, charges: str | None = None) -> "Mesh":
        """Surface from a .off or .gts file, optional 'x y z q [r]' charges."""
        h = vp()
        check_host(lib().beep_mesh_read(str(path).encode(), C.byref(h)))
        m = cls(h)
        if charges is not None:
            check_host(lib().beep_mesh_read_charges(m._h, str(charges).encode()))
        return m

    @classmethod
    def read_mtz(cls, path: str) -> "Mesh":
        """A .mtz mesh tarball: surface, charges and centre from the members
        its definition.xml names."""
        h = vp()
        check_host(lib().beep_mesh_read_mtz(str(path).encode(), C.byref(h)))
        return cls(h)

    def initial_fh(self):
        """(f, h) per node patch from a .mtz fh member, or None."""
        n = self.info()["n_vertices"]
        f, h = np.empty(n), np.empty(n)
        dp = C.POINTER(C.c_double)
        rc = lib().beep_mesh_initial_fh(self._h, f.ctypes.data_as(dp), h.ctypes.data_as(dp))
        if rc < 0:
            check_host(rc)
        return None if rc == 1 else (f, h)

    def set_curved(self, on: bool = True) -> "Mesh":
        """Curved (PNG1) node patches instead of planar ones."""
        check_host(lib().beep_mesh_set_curved(self._h, 1 if on else 0))
        self.curved = bool(on)
        return self

    def curved_point(self, tri: int, u: float, v: float):
        """(point, unit normal) of curved triangle `tri` at barycentric (u, v)."""
        p, n = np.zeros(3), np.zeros(3)
        check_host(lib().beep_mesh_curved_point(self._h, int(tri), float(u), float(v), _dp(p), _dp(n)))
        return p, n

    def clone(self) -> "Mesh":
        h = vp()
        check_host(lib().beep_mesh_clone(self._h, C.byref(h)))
        return Mesh(h)

    def write_off(self, path: str):
        check_host(lib().beep_mesh_write_off(self._h, str(path).encode()))

    def set_charges(self, xyzq) -> "Mesh":
        a = np.ascontiguousarray(xyzq, dtype=np.float64).reshape(-1, 4)
        check_host(lib().beep_mesh_set_charges(self._h, a.shape[0], _dp(a)))
        return self

    def charges_ball(self, n, rmax, qmin, qmax, seed) -> "Mesh":
        check_host(lib().beep_mesh_charges_ball(self._h, n, rmax, qmin, qmax, seed))
        return self

    def charges_walk(self, n, step, clearance, qsigma, seed) -> "Mesh":
        check_host(lib().beep_mesh_charges_walk(self._h, n, step, clearance, qsigma, seed))
        return self

    def info(self):
        nv, nt, nc = C.c_int(), C.c_int(), C.c_int()
        c = np.zeros(3)
        r = C.c_double()
        check_host(lib().beep_mesh_info(self._h, C.byref(nv), C.byref(nt), C.byref(nc), _dp(c),
                                        C.byref(r)))
        return {"n_vertices": nv.value, "n_triangles": nt.value, "n_charges": nc.value,
                "centre": c, "radius": r.value}

    def vertices(self):
        i = self.info()
        xyz = np.zeros((i["n_vertices"], 3))
        tri = np.zeros((i["n_triangles"], 3), dtype=np.int32)
        check_host(lib().beep_mesh_vertices(self._h, _dp(xyz), _ip(tri)))
        return xyz, tri

    def __del__(self):
        try:
            if self._h:
                lib().beep_mesh_destroy(self._h)
        except Exception:
            pass


class Scene:
    """Solver parameters + rigid mesh instances + octree."""

    def __init__(self, Dsolvent=80.0, kappa=0.0, quad=0, qual=4, nbsize=1500, host_points=True,
                 tree=True):
        h = vp()
        check_host(lib().beep_scene_create(Dsolvent, kappa, quad, qual, nbsize, C.byref(h)))
        self._h = h
        self.host_points = bool(host_points)
        if not self.host_points:
            check_host(lib().beep_scene_set_host_points(h, 0))
        # tree=False: patches only (the device builds the octree, Device.build_tree)
        if not tree:
            check_host(lib().beep_scene_set_tree(h, 0))
        self.Dsolvent, self.kappa, self.quad, self.qual, self.nbsize = Dsolvent, kappa, quad, qual, nbsize
        self._meshes = []
        self._arrays = None

    def add_instance(self, mesh: Mesh, Dprotein: float, quat=None, offset=None) -> "Scene":
        q = None if quat is None else np.ascontiguousarray(quat, dtype=np.float64)
        o = None if offset is None else np.ascontiguousarray(offset, dtype=np.float64)
        check_host(lib().beep_scene_add_instance(self._h, mesh._h, None if q is None else _dp(q),
                                                 None if o is None else _dp(o), Dprotein))
        self._meshes.append(mesh)
        self._arrays = None
        return self

    def build(self) -> "Scene":
        check_host(lib().beep_scene_build(self._h))
        self._arrays = None
        return self

    def sizes(self) -> SceneSizes:
        s = SceneSizes()
        check_host(lib().beep_scene_get_sizes(self._h, C.byref(s)))
        return s

    def arrays(self) -> dict:
        """All flat arrays of the scene (geometry, point sets, tree, charges)."""
        if self._arrays is not None:
            return self._arrays
        s = self.sizes()
        N = s.n_patches
        a = {
            "centroid": np.zeros((N, 3)), "normal": np.zeros((N, 3)), "area": np.zeros(N),
            "eps": np.zeros(N), "gc": np.zeros(N), "instance": np.zeros(N, dtype=np.int32),
        }
        check_host(lib().beep_scene_patches(self._h, _dp(a["centroid"]), _dp(a["normal"]),
                                            _dp(a["area"]), _dp(a["eps"]), _dp(a["gc"]),
                                            _ip(a["instance"])))
        for kind, name, n in ((0, "qual", s.n_qual), (1, "quad", s.n_quad), (2, "gal", s.n_galerkin)):
            ptr = np.zeros(N + 1, dtype=np.int32)
            pts = np.zeros((n if self.host_points else 0, 7))
            check_host(lib().beep_scene_points(self._h, kind, _ip(ptr),
                                               _dp(pts) if self.host_points else None))
            a[name + "_ptr"], a[name] = ptr, pts
        L = s.n_leaves
        a["leaf_ptr"] = np.zeros(L + 1, dtype=np.int32)
        a["leaf_idx"] = np.zeros(N, dtype=np.int32)
        a["nb_ptr"] = np.zeros(L + 1, dtype=np.int32)
        a["nb_idx"] = np.zeros(s.n_near_cols, dtype=np.int32)
        check_host(lib().beep_scene_tree(self._h, _ip(a["leaf_ptr"]), _ip(a["leaf_idx"]),
                                         _ip(a["nb_ptr"]), _ip(a["nb_idx"])))
        a["charges"] = np.zeros((s.n_charges, 4))
        a["charge_instance"] = np.zeros(s.n_charges, dtype=np.int32)
        check_host(lib().beep_scene_charges(self._h, _dp(a["charges"]), _ip(a["charge_instance"])))
        a["Dprotein"] = np.zeros(s.n_instances)
        a["inst_patch_ptr"] = np.zeros(s.n_instances + 1, dtype=np.int32)
        a["inst_charge_ptr"] = np.zeros(s.n_instances + 1, dtype=np.int32)
        check_host(lib().beep_scene_instances(self._h, _dp(a["Dprotein"]), _ip(a["inst_patch_ptr"]),
                                              _ip(a["inst_charge_ptr"])))
        self._arrays = a
        return a

    def surface(self) -> dict:
        """Placed vertices, triangles and the triangles around every vertex
        (with the vertex's corner in each), as the device generates points;
        for curved meshes also each triangle's index among the curved ones
        (-1: planar) and their placed control points (n_curved x 63)."""
        s = self.sizes()
        out = {"verts": np.zeros((s.n_patches, 3)), "tris": np.zeros((s.n_triangles, 3), np.int32),
               "vt_ptr": np.zeros(s.n_patches + 1, np.int32),
               "vt_tri": np.zeros(s.n_vertex_triangles, np.int32),
               "vt_corner": np.zeros(s.n_vertex_triangles, np.int32)}
        check_host(lib().beep_scene_surface(self._h, _dp(out["verts"]), _ip(out["tris"]),
                                            _ip(out["vt_ptr"]), _ip(out["vt_tri"]),
                                            _ip(out["vt_corner"])))
        nc = C.c_int(0)
        check_host(lib().beep_scene_surface_curved(self._h, C.byref(nc), None, None))
        out["ctl_idx"] = np.zeros(s.n_triangles, np.int32)
        out["ctl"] = np.zeros((nc.value, 63))
        check_host(lib().beep_scene_surface_curved(self._h, None, _ip(out["ctl_idx"]), _dp(out["ctl"])))
        return out

    def __del__(self):
        try:
            if self._h:
                lib().beep_scene_destroy(self._h)
        except Exception:
            pass
