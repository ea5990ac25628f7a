"""Python view of the device boundary (include/beep_hip.h)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import (BeepError, FarInfo, HipGeometry, HipSurface, HipTree, MatvecTiming,
                   SolveOpts, SolveStats, check_hip, dp, ip, lib, vp)

FP64, FP32, FP32_RESID = 0, 1, 2
SOLVE_FP64, SOLVE_MIXED, SOLVE_FP32 = 0, 1, 2
PRECISIONS = {"fp64": FP64, "fp32": FP32, "fp32_resid": FP32_RESID}
SOLVE_MODES = {"fp64": SOLVE_FP64, "mixed": SOLVE_MIXED, "fp32": SOLVE_FP32}


def _dp(a):
    return a.ctypes.data_as(dp)


def _ip(a):
    return a.ctypes.data_as(ip)


def _prec(p):
    return PRECISIONS[p] if isinstance(p, str) else int(p)


class Device:
    """One HIP context on one GPU (beep_hip_create)."""

    def __init__(self, device: int = 0):
        L = lib()
        if not hasattr(L, "beep_hip_create"):
            raise BeepError("libbeep3.so was built without the HIP boundary")
        h = vp()
        rc = L.beep_hip_create(device, C.byref(h))
        if rc != 0:
            raise BeepError(L.beep_hip_last_error(None).decode())
        self._h = h
        self.n = 0
        self._keep = []

    def _check(self, rc):
        check_hip(self._h, rc)

    def set_geometry(self, a: dict):
        self.n = n = int(a["area"].shape[0])
        keep = {k: np.ascontiguousarray(a[k]) for k in (
            "centroid", "normal", "area", "eps", "gc", "qual_ptr", "qual", "quad_ptr", "quad",
            "gal_ptr", "gal")}
        g = HipGeometry()
        g.n_patches = n
        g.centroid, g.normal = keep["centroid"].ctypes.data, keep["normal"].ctypes.data
        g.area, g.eps_ratio, g.gc = keep["area"].ctypes.data, keep["eps"].ctypes.data, keep["gc"].ctypes.data
        g.qual_ptr, g.qual_pts = keep["qual_ptr"].ctypes.data, keep["qual"].ctypes.data
        g.quad_ptr, g.quad_pts = keep["quad_ptr"].ctypes.data, keep["quad"].ctypes.data
        g.gal_ptr, g.gal_pts = keep["gal_ptr"].ctypes.data, keep["gal"].ctypes.data
        self._check(lib().beep_hip_set_geometry(self._h, C.byref(g)))

    def set_surface(self, a: dict, surf: dict, qual_rule: int, quad_rule: int):
        """Geometry from the surface: the device generates the point sets
        (a: patch arrays of Scene.arrays(), surf: Scene.surface())."""
        self.n = n = int(a["area"].shape[0])
        keep = {k: np.ascontiguousarray(a[k]) for k in ("centroid", "normal", "area", "eps", "gc")}
        keep.update({k: np.ascontiguousarray(surf[k]) for k in surf})
        g = HipSurface()
        g.n_patches = n
        g.centroid, g.normal = keep["centroid"].ctypes.data, keep["normal"].ctypes.data
        g.area, g.eps_ratio, g.gc = keep["area"].ctypes.data, keep["eps"].ctypes.data, keep["gc"].ctypes.data
        g.verts, g.n_tris, g.tris = keep["verts"].ctypes.data, int(keep["tris"].shape[0]), keep["tris"].ctypes.data
        g.vt_ptr, g.vt_tri, g.vt_corner = (keep["vt_ptr"].ctypes.data, keep["vt_tri"].ctypes.data,
                                           keep["vt_corner"].ctypes.data)
        g.qual_rule, g.quad_rule = int(qual_rule), int(quad_rule)
        if "ctl" in keep and keep["ctl"].shape[0] > 0:
            g.n_curved = int(keep["ctl"].shape[0])
            g.ctl_idx, g.ctl = keep["ctl_idx"].ctypes.data, keep["ctl"].ctypes.data
        self._check(lib().beep_hip_set_surface(self._h, C.byref(g)))

    def set_tree(self, a: dict):
        keep = {k: np.ascontiguousarray(a[k], dtype=np.int32) for k in (
            "leaf_ptr", "leaf_idx", "nb_ptr", "nb_idx")}
        t = HipTree()
        t.n_leaves = int(keep["leaf_ptr"].shape[0] - 1)
        t.leaf_ptr, t.leaf_idx = keep["leaf_ptr"].ctypes.data, keep["leaf_idx"].ctypes.data
        t.nb_ptr, t.nb_idx = keep["nb_ptr"].ctypes.data, keep["nb_idx"].ctypes.data
        self._check(lib().beep_hip_set_tree(self._h, C.byref(t)))

    def build_tree(self, centre, edge: float, nbsize: int):
        """The octree and near list built on the device from the centroids of
        the geometry set (the host's octree, same leaves and lists), in the
        bounding cube (centre, edge).  Returns (leaves, neighbourhood entries)."""
        c = np.ascontiguousarray(centre, dtype=np.float64).reshape(3)
        nl, nn = C.c_int(0), C.c_longlong(0)
        self._check(lib().beep_hip_build_tree(self._h, c.ctypes.data_as(C.POINTER(C.c_double)),
                                              float(edge), int(nbsize), C.byref(nl), C.byref(nn)))
        self._tree_sizes = (nl.value, nn.value)
        return self._tree_sizes

    def tree_arrays(self, lists: bool = True) -> dict:
        """leaf_ptr, leaf_idx, nb_ptr, nb_idx of the tree last built on the
        device (lists=False: the two pointer arrays only)."""
        L, nnb = self._tree_sizes
        out = {"leaf_ptr": np.zeros(L + 1, np.int32), "nb_ptr": np.zeros(L + 1, np.int32)}
        if lists:
            out["leaf_idx"] = np.zeros(self.n, np.int32)
            out["nb_idx"] = np.zeros(max(nnb, 1), np.int32)
        ip_ = C.POINTER(C.c_int)
        ptr = {k: (out[k].ctypes.data_as(ip_) if k in out else None) for k in (
            "leaf_ptr", "leaf_idx", "nb_ptr", "nb_idx")}
        self._check(lib().beep_hip_get_tree(self._h, ptr["leaf_ptr"], ptr["leaf_idx"], ptr["nb_ptr"],
                                            ptr["nb_idx"]))
        if lists:
            out["nb_idx"] = out["nb_idx"][:nnb]
        return out

    FAR_MODES = {"auto": -1, "direct": 0, "fmm": 1, "pw": 2}

    def set_far_field(self, mode="auto", order=4, leaf_target=48):
        m = self.FAR_MODES[mode] if isinstance(mode, str) else int(mode)
        self._check(lib().beep_hip_set_far_field(self._h, m, int(order), int(leaf_target)))

    def set_overlap(self, on: bool = True):
        """Far field on a second stream, concurrent with the near SpMV."""
        self._check(lib().beep_hip_set_overlap(self._h, 1 if on else 0))

    def set_graphs(self, on: bool = True):
        """GMRES Arnoldi steps as one replayed HIP graph (default) or kernel
        by kernel."""
        self._check(lib().beep_hip_set_graphs(self._h, 1 if on else 0))

    def set_near_singular_fp64(self, on: bool = True):
        """FP64 integration of the near-singular pairs of the FP32 set
        (default), from the next FP32 precompute."""
        self._check(lib().beep_hip_set_near_singular_fp64(self._h, 1 if on else 0))

    NEAR_MODES = {"precompute": 0, "on_the_fly": 1}

    def set_near_mode(self, mode="precompute"):
        """"precompute" (stored coefficients, the default) or "on_the_fly"
        (every mat-vec integrates the explicit pairs; nothing stored).  Drops
        the precomputed sets: precompute again."""
        m = self.NEAR_MODES[mode] if isinstance(mode, str) else int(mode)
        self._check(lib().beep_hip_set_near_mode(self._h, m))

    def share_near(self, src: "Device", copy_offsets, pair_i=None, pair_j=None):
        """Near field of src's coefficient sets applied to copies of its
        patches at the given offsets, plus the listed explicit pairs."""
        off = np.ascontiguousarray(copy_offsets, dtype=np.int32)
        pi = np.ascontiguousarray([] if pair_i is None else pair_i, dtype=np.int32)
        pj = np.ascontiguousarray([] if pair_j is None else pair_j, dtype=np.int32)
        if pi.shape != pj.shape:
            raise ValueError("pair_i and pair_j differ in length")
        self._share_src = src  # src must outlive this context's use of it
        self._check(lib().beep_hip_share_near(self._h, src._h, len(off), _ip(off), len(pi),
                                              _ip(pi) if len(pi) else None, _ip(pj) if len(pj) else None))

    def share_near_cutoff(self, src: "Device", copy_offsets, cutoff: float) -> int:
        """As share_near, with the pairs between copies closer than cutoff
        found on the device; returns their count."""
        off = np.ascontiguousarray(copy_offsets, dtype=np.int32)
        n = C.c_longlong()
        self._share_src = src
        self._check(lib().beep_hip_share_near_cutoff(self._h, src._h, len(off), _ip(off), float(cutoff),
                                                     C.byref(n)))
        return n.value

    def set_gram_schmidt(self, passes: int = 1):
        """Gram-Schmidt passes per Arnoldi step: 1 (classical, the
        reference's) or 2 (CGS2)."""
        self._check(lib().beep_hip_set_gram_schmidt(self._h, int(passes)))

    def set_guard(self, max_value: float = 1e6):
        self._check(lib().beep_hip_set_guard(self._h, float(max_value)))

    def guard_stats(self) -> dict:
        """Entries repaired by the last precompute of each set and by the
        last sanity check."""
        a = (C.c_longlong * 4)()
        self._check(lib().beep_hip_guard_stats(self._h, a))
        return {"fp64": a[0], "fp32": a[1], "diag": a[2], "sanity_check": a[3]}

    def sanity_check(self) -> int:
        """BEEP::sanity_check: repair stored coefficients that are out of
        range or not finite; returns how many were repaired."""
        n = C.c_longlong()
        self._check(lib().beep_hip_sanity_check(self._h, C.byref(n)))
        return n.value

    def debug_poke(self, which: str, entry: int, values):
        """Test hook: overwrite one stored entry ('fp64', 'fp32' or 'diag')."""
        v = np.ascontiguousarray(values, dtype=np.float64).reshape(4)
        w = {"fp64": 0, "fp32": 1, "diag": 2}[which]
        self._check(lib().beep_hip_debug_poke(self._h, w, int(entry), _dp(v)))

    def near_singular_count(self):
        """(entries computed in FP64, entries left in FP32 for lack of room)
        of the current FP32 set."""
        a, b = C.c_longlong(), C.c_longlong()
        self._check(lib().beep_hip_near_singular_count(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def autotune_leaf(self, precision="fp64", candidates=(24, 32, 40, 64, 96, 160), reps=3):
        """Time mat-vecs for each FMM leaf target, keep the fastest; returns
        (best, {leaf: ms per mat-vec})."""
        c = np.ascontiguousarray(candidates, dtype=np.int32)
        ms = np.zeros(len(c))
        best = C.c_int()
        self._check(lib().beep_hip_autotune_leaf(
            self._h, _prec(precision), c.ctypes.data_as(C.POINTER(C.c_int)), len(c), int(reps),
            ms.ctypes.data_as(C.POINTER(C.c_double)), C.byref(best)))
        return best.value, {int(k): float(v) for k, v in zip(c, ms)}

    def autotune_spmv(self, precision="fp64", candidates=(0, 1, 2, 3, 4, 6), reps=3):
        """Time whole mat-vecs for each near-SpMV residency cap (blocks per
        CU, 0 = uncapped), keep the fastest; returns (best, {cap: ms})."""
        c = np.ascontiguousarray(candidates, dtype=np.int32)
        ms = np.zeros(len(c))
        best = C.c_int()
        self._check(lib().beep_hip_autotune_spmv(
            self._h, _prec(precision), c.ctypes.data_as(C.POINTER(C.c_int)), len(c), int(reps),
            ms.ctypes.data_as(C.POINTER(C.c_double)), C.byref(best)))
        return best.value, {int(k): float(v) for k, v in zip(c, ms)}

    def set_spmv_residency(self, blocks_per_cu: int):
        self._check(lib().beep_hip_set_spmv_residency(self._h, int(blocks_per_cu)))

    def autotune_spmv_depth(self, precision="fp64", candidates=(1, 2, 4), reps=3):
        """Time whole mat-vecs for each near-SpMV load depth, keep the
        fastest; returns (best, {depth: ms})."""
        c = np.ascontiguousarray(candidates, dtype=np.int32)
        ms = np.zeros(len(c))
        best = C.c_int()
        self._check(lib().beep_hip_autotune_spmv_depth(
            self._h, _prec(precision), c.ctypes.data_as(C.POINTER(C.c_int)), len(c), int(reps),
            ms.ctypes.data_as(C.POINTER(C.c_double)), C.byref(best)))
        return best.value, {int(k): float(v) for k, v in zip(c, ms)}

    def set_spmv_depth(self, depth: int):
        self._check(lib().beep_hip_set_spmv_depth(self._h, int(depth)))

    def far_info(self) -> dict:
        f = FarInfo()
        self._check(lib().beep_hip_far_info(self._h, C.byref(f)))
        return {k: getattr(f, k) for k, _ in FarInfo._fields_}

    def precompute(self, kappa: float, precision="fp32"):
        self._check(lib().beep_hip_precompute(self._h, float(kappa), _prec(precision)))

    def matvec(self, x, precision="fp64"):
        x = np.ascontiguousarray(x, dtype=np.float64)
        if x.shape != (2 * self.n,):
            raise ValueError(f"x must have shape ({2 * self.n},)")
        y = np.empty_like(x)
        self._check(lib().beep_hip_matvec(self._h, _dp(x), _dp(y), _prec(precision)))
        return y

    def matvec_device(self, d_x: int, d_y: int, precision="fp64"):
        self._check(lib().beep_hip_matvec_device(self._h, vp(d_x), vp(d_y), _prec(precision)))

    def timing(self) -> MatvecTiming:
        t = MatvecTiming()
        self._check(lib().beep_hip_matvec_timing(self._h, C.byref(t)))
        return t

    def timing_reset(self):
        self._check(lib().beep_hip_timing_reset(self._h))

    def rhs(self, xyzq, Dsolvent: float):
        q = np.ascontiguousarray(xyzq, dtype=np.float64).reshape(-1, 4)
        b = np.zeros(2 * self.n)
        self._check(lib().beep_hip_rhs(self._h, q.shape[0], _dp(q), float(Dsolvent), _dp(b)))
        return b

    def energy_coefficients(self, inst_patch_ptr, inst_charge_ptr, xyzq):
        pp = np.ascontiguousarray(inst_patch_ptr, dtype=np.int32)
        cp = np.ascontiguousarray(inst_charge_ptr, dtype=np.int32)
        q = np.ascontiguousarray(xyzq, dtype=np.float64).reshape(-1, 4)
        ef, eh = np.zeros(self.n), np.zeros(self.n)
        self._check(lib().beep_hip_energy_coefficients(self._h, len(pp) - 1, _ip(pp), _ip(cp),
                                                       _dp(q), _dp(ef), _dp(eh)))
        return ef, eh

    def solve(self, b, x0=None, tol=1e-6, max_iter=100, mode="fp64", max_refine=10,
              inner_tol=0.0, verbose=False, resid_fp64=False):
        """inner_tol: tolerance of the inner FP32 solves of the mixed mode
        (0: half of tol); resid_fp64: the mixed mode's residual with the FP64
        coefficient set (which must be precomputed) instead of the FP32 one."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=np.float64)
        o = SolveOpts(tol, max_iter, SOLVE_MODES[mode] if isinstance(mode, str) else int(mode),
                      max_refine, inner_tol, int(verbose), int(resid_fp64))
        st = SolveStats()
        self._check(lib().beep_hip_solve(self._h, _dp(b), _dp(x), C.byref(o), C.byref(st)))
        return x, {f: getattr(st, f) for f, _ in SolveStats._fields_}

    def rhs_device(self, xyzq, Dsolvent: float, d_b: int):
        """b = [f_rhs; h_rhs] into the device vector d_b (2N doubles)."""
        q = np.ascontiguousarray(xyzq, dtype=np.float64).reshape(-1, 4)
        self._check(lib().beep_hip_rhs_device(self._h, q.shape[0], _dp(q), float(Dsolvent),
                                              vp(d_b)))

    def solve_device(self, d_b: int, d_x: int, tol=1e-6, max_iter=100, mode="fp64",
                     max_refine=10, inner_tol=0.0, verbose=False, resid_fp64=False):
        """solve on device vectors (d_x: initial guess in, solution out)."""
        o = SolveOpts(tol, max_iter, SOLVE_MODES[mode] if isinstance(mode, str) else int(mode),
                      max_refine, inner_tol, int(verbose), int(resid_fp64))
        st = SolveStats()
        self._check(lib().beep_hip_solve_device(self._h, vp(d_b), vp(d_x), C.byref(o),
                                                C.byref(st)))
        return {f: getattr(st, f) for f, _ in SolveStats._fields_}

    # device memory helpers (bench / serving keep vectors resident)
    def alloc(self, nbytes: int) -> int:
        p = vp()
        self._check(lib().beep_hip_device_alloc(self._h, nbytes, C.byref(p)))
        return p.value

    def free(self, p: int):
        self._check(lib().beep_hip_device_free(self._h, vp(p)))

    def upload(self, d_ptr: int, arr):
        a = np.ascontiguousarray(arr)
        self._check(lib().beep_hip_memcpy(self._h, vp(d_ptr), a.ctypes.data_as(vp), a.nbytes, 1))

    def download(self, arr, d_ptr: int):
        self._check(lib().beep_hip_memcpy(self._h, arr.ctypes.data_as(vp), vp(d_ptr), arr.nbytes, 2))
        return arr

    def copy_device(self, d_dst: int, d_src: int, nbytes: int):
        self._check(lib().beep_hip_memcpy(self._h, vp(d_dst), vp(d_src), int(nbytes), 3))

    def gather_rows64(self, d_src: int, d_idx: int, rows: int, d_dst: int):
        """dst row r = src row idx[r] over 64-byte rows (a measurement aid)."""
        self._check(lib().beep_hip_gather_rows64(self._h, vp(d_src), vp(d_idx), int(rows), vp(d_dst)))

    def synchronize(self):
        self._check(lib().beep_hip_synchronize(self._h))

    def memory_info(self):
        f, t = C.c_size_t(), C.c_size_t()
        self._check(lib().beep_hip_memory_info(self._h, C.byref(f), C.byref(t)))
        return f.value, t.value

    def memory_held(self) -> int:
        """Device bytes held by this context's buffers."""
        b = C.c_size_t()
        self._check(lib().beep_hip_memory_held(self._h, C.byref(b)))
        return b.value

    def close(self):
        if getattr(self, "_h", None):
            lib().beep_hip_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
