"""Loader for the in-tree native library ``libbeep3.so``.

The library holds the host-side scene code (include/beep_host.h) and the HIP
device boundary (include/beep_hip.h).  It is built by ``beep3_amd.build`` (or
``__graft_entry__.build()``); a missing library is an error, never a silent
fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbeep3.so")

_lib = None


class BeepError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise BeepError(
                f"{LIB_PATH} is missing: build it with `python -m beep3_amd.build` "
                "(the HIP extension is required; there is no CPU fallback)")
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)
vp = C.c_void_p


class SceneSizes(C.Structure):
    _fields_ = [
        ("n_patches", C.c_int), ("n_instances", C.c_int), ("n_charges", C.c_int),
        ("n_qual", C.c_int), ("n_quad", C.c_int), ("n_galerkin", C.c_int),
        ("n_leaves", C.c_int), ("n_near_cols", C.c_int), ("n_near_pairs", C.c_longlong),
        ("tree_depth", C.c_int), ("cube_centre", C.c_double * 3), ("cube_edge", C.c_double),
        ("n_triangles", C.c_int), ("n_vertex_triangles", C.c_int),
    ]


class HipGeometry(C.Structure):
    _fields_ = [
        ("n_patches", C.c_int),
        ("centroid", vp), ("normal", vp), ("area", vp), ("eps_ratio", vp), ("gc", vp),
        ("qual_ptr", vp), ("qual_pts", vp),
        ("quad_ptr", vp), ("quad_pts", vp),
        ("gal_ptr", vp), ("gal_pts", vp),
    ]


class HipSurface(C.Structure):
    _fields_ = [
        ("n_patches", C.c_int),
        ("centroid", vp), ("normal", vp), ("area", vp), ("eps_ratio", vp), ("gc", vp),
        ("verts", vp), ("n_tris", C.c_int), ("tris", vp), ("vt_ptr", vp), ("vt_tri", vp),
        ("vt_corner", vp), ("qual_rule", C.c_int), ("quad_rule", C.c_int),
        ("n_curved", C.c_int), ("ctl_idx", vp), ("ctl", vp),
    ]


class HipTree(C.Structure):
    _fields_ = [
        ("n_leaves", C.c_int),
        ("leaf_ptr", vp), ("leaf_idx", vp), ("nb_ptr", vp), ("nb_idx", vp),
    ]


class SolveOpts(C.Structure):
    _fields_ = [
        ("tol", C.c_double), ("max_iter", C.c_int), ("mode", C.c_int),
        ("max_refine", C.c_int), ("inner_tol", C.c_double), ("verbose", C.c_int),
        ("resid_fp64", C.c_int),
    ]


class SolveStats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int), ("matvecs", C.c_int), ("refinements", C.c_int),
        ("converged", C.c_int), ("residual", C.c_double), ("rel_residual", C.c_double),
        ("time_ms", C.c_double), ("matvec_ms", C.c_double),
    ]


class MatvecTiming(C.Structure):
    _fields_ = [
        ("near_ms", C.c_double), ("far_ms", C.c_double), ("total_ms", C.c_double),
        ("count", C.c_int), ("near_ms_sum", C.c_double), ("far_ms_sum", C.c_double),
        ("total_ms_sum", C.c_double), ("m2l_ms", C.c_double), ("m2l_ms_sum", C.c_double),
    ]


class FarInfo(C.Structure):
    _fields_ = [
        ("mode", C.c_int), ("order", C.c_int), ("leaf_level", C.c_int), ("n_leaf_boxes", C.c_int),
        ("n_boxes", C.c_int), ("m2l_pairs", C.c_longlong), ("p2p_pairs", C.c_longlong),
        ("p2p_residual_pairs", C.c_longlong), ("plan_ms", C.c_double),
        ("bytes_fp64", C.c_double), ("bytes_fp32", C.c_double),
    ]


class ServerOpts(C.Structure):
    _fields_ = [
        ("device", C.c_int), ("slots", C.c_int), ("hbm_budget", C.c_double),
        ("solve_mode", C.c_int), ("tol", C.c_double), ("max_iter", C.c_int),
        ("max_refine", C.c_int), ("inner_tol", C.c_double), ("far_mode", C.c_int),
        ("fmm_order", C.c_int), ("fmm_leaf", C.c_int), ("keep_fh", C.c_int),
        ("near_mode", C.c_int), ("device_tree", C.c_int),
    ]


class ServerResult(C.Structure):
    _fields_ = [
        ("id", C.c_longlong), ("status", C.c_int), ("n_patches", C.c_int),
        ("energy", C.c_double), ("iterations", C.c_int), ("matvecs", C.c_int),
        ("refinements", C.c_int), ("converged", C.c_int), ("rel_residual", C.c_double),
        ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double),
        ("host_ms", C.c_double), ("device_ms", C.c_double), ("upload_ms", C.c_double),
        ("precompute_ms", C.c_double), ("solve_ms", C.c_double), ("bytes", C.c_double),
        ("device_free", C.c_double), ("slot", C.c_int), ("message", C.c_char * 128),
        ("near_on_the_fly", C.c_int),
    ]


def _declare(L):
    L.beep_host_last_error.restype = C.c_char_p
    L.beep_mesh_icosphere.argtypes = [C.c_int, C.c_double, C.POINTER(vp)]
    L.beep_mesh_star.argtypes = [C.c_int, C.c_double, C.c_int, C.c_double, C.c_uint64, C.POINTER(vp)]
    L.beep_mesh_set_charges.argtypes = [vp, C.c_int, dp]
    L.beep_mesh_charges_ball.argtypes = [vp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_uint64]
    L.beep_mesh_charges_walk.argtypes = [vp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_uint64]
    L.beep_mesh_info.argtypes = [vp, ip, ip, ip, dp, dp]
    L.beep_mesh_vertices.argtypes = [vp, dp, ip]
    L.beep_mesh_read.argtypes = [C.c_char_p, C.POINTER(vp)]
    L.beep_mesh_read_mtz.argtypes = [C.c_char_p, C.POINTER(vp)]
    L.beep_mesh_set_curved.argtypes = [vp, C.c_int]
    L.beep_mesh_curved_point.argtypes = [vp, C.c_int, C.c_double, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.beep_mesh_initial_fh.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.beep_mesh_write_off.argtypes = [vp, C.c_char_p]
    L.beep_mesh_read_charges.argtypes = [vp, C.c_char_p]
    L.beep_mesh_clone.argtypes = [vp, C.POINTER(vp)]
    L.beep_mesh_destroy.argtypes = [vp]
    L.beep_mesh_destroy.restype = None
    L.beep_scene_create.argtypes = [C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, C.POINTER(vp)]
    L.beep_scene_add_instance.argtypes = [vp, vp, dp, dp, C.c_double]
    L.beep_scene_build.argtypes = [vp]
    L.beep_scene_get_sizes.argtypes = [vp, C.POINTER(SceneSizes)]
    L.beep_scene_patches.argtypes = [vp, dp, dp, dp, dp, dp, ip]
    L.beep_scene_points.argtypes = [vp, C.c_int, ip, dp]
    L.beep_scene_tree.argtypes = [vp, ip, ip, ip, ip]
    L.beep_scene_charges.argtypes = [vp, dp, ip]
    L.beep_scene_instances.argtypes = [vp, dp, ip, ip]
    L.beep_scene_destroy.argtypes = [vp]
    L.beep_scene_destroy.restype = None
    L.beep_scene_set_host_points.argtypes = [vp, C.c_int]
    L.beep_scene_set_tree.argtypes = [vp, C.c_int]
    L.beep_scene_surface.argtypes = [vp, dp, ip, ip, ip, ip]
    L.beep_scene_surface_curved.argtypes = [vp, ip, ip, dp]
    L.beep_server_default_opts.argtypes = [C.POINTER(ServerOpts)]
    L.beep_server_default_opts.restype = None
    L.beep_server_estimate.argtypes = [C.POINTER(ServerOpts), vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.beep_server_create.argtypes = [C.POINTER(ServerOpts), C.POINTER(vp)]
    L.beep_server_submit.argtypes = [vp, vp, C.POINTER(C.c_longlong)]
    L.beep_server_poll.argtypes = [vp, C.c_double, C.POINTER(ServerResult)]
    L.beep_server_take_fh.argtypes = [vp, C.c_longlong, dp, dp]
    L.beep_server_clock.argtypes = [vp]
    L.beep_server_clock.restype = C.c_double
    L.beep_server_pending.argtypes = [vp]
    L.beep_server_last_error.argtypes = [vp]
    L.beep_server_last_error.restype = C.c_char_p
    L.beep_server_destroy.argtypes = [vp]
    _declare_poses(L)
    if hasattr(L, "beep_hip_create"):
        _declare_hip(L)


def _declare_hip(L):
    L.beep_hip_last_error.restype = C.c_char_p
    L.beep_hip_last_error.argtypes = [vp]
    L.beep_hip_create.argtypes = [C.c_int, C.POINTER(vp)]
    L.beep_hip_destroy.argtypes = [vp]
    L.beep_hip_set_geometry.argtypes = [vp, C.POINTER(HipGeometry)]
    L.beep_hip_set_tree.argtypes = [vp, C.POINTER(HipTree)]
    L.beep_hip_build_tree.argtypes = [vp, dp, C.c_double, C.c_int, ip, C.POINTER(C.c_longlong)]
    L.beep_hip_get_tree.argtypes = [vp, ip, ip, ip, ip]
    L.beep_hip_set_surface.argtypes = [vp, C.POINTER(HipSurface)]
    L.beep_hip_precompute.argtypes = [vp, C.c_double, C.c_int]
    L.beep_hip_set_far_field.argtypes = [vp, C.c_int, C.c_int, C.c_int]
    L.beep_hip_set_overlap.argtypes = [vp, C.c_int]
    L.beep_hip_set_graphs.argtypes = [vp, C.c_int]
    L.beep_hip_set_near_singular_fp64.argtypes = [vp, C.c_int]
    L.beep_hip_set_near_mode.argtypes = [vp, C.c_int]
    L.beep_hip_set_guard.argtypes = [vp, C.c_double]
    L.beep_hip_share_near.argtypes = [vp, vp, C.c_int, ip, C.c_longlong, ip, ip]
    L.beep_hip_share_near_cutoff.argtypes = [vp, vp, C.c_int, ip, C.c_double, C.POINTER(C.c_longlong)]
    L.beep_hip_set_gram_schmidt.argtypes = [vp, C.c_int]
    L.beep_hip_guard_stats.argtypes = [vp, C.POINTER(C.c_longlong)]
    L.beep_hip_sanity_check.argtypes = [vp, C.POINTER(C.c_longlong)]
    L.beep_hip_debug_poke.argtypes = [vp, C.c_int, C.c_longlong, dp]
    L.beep_hip_near_singular_count.argtypes = [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
    L.beep_hip_autotune_leaf.argtypes = [vp, C.c_int, ip, C.c_int, C.c_int, dp, ip]
    L.beep_hip_autotune_spmv.argtypes = [vp, C.c_int, ip, C.c_int, C.c_int, dp, ip]
    L.beep_hip_set_spmv_residency.argtypes = [vp, C.c_int]
    L.beep_hip_autotune_spmv_depth.argtypes = [vp, C.c_int, ip, C.c_int, C.c_int, dp, ip]
    L.beep_hip_set_spmv_depth.argtypes = [vp, C.c_int]
    L.beep_hip_far_info.argtypes = [vp, C.POINTER(FarInfo)]
    L.beep_hip_matvec.argtypes = [vp, dp, dp, C.c_int]
    L.beep_hip_matvec_device.argtypes = [vp, vp, vp, C.c_int]
    L.beep_hip_matvec_device_noalloc.argtypes = [vp, vp, vp, C.c_int]
    L.beep_hip_pw_selftest.argtypes = [C.c_int, C.c_int, C.c_double, dp]
    L.beep_hip_matvec_timing.argtypes = [vp, C.POINTER(MatvecTiming)]
    L.beep_hip_timing_reset.argtypes = [vp]
    L.beep_hip_rhs.argtypes = [vp, C.c_int, dp, C.c_double, dp]
    L.beep_hip_energy_coefficients.argtypes = [vp, C.c_int, ip, ip, dp, dp, dp]
    L.beep_hip_solve.argtypes = [vp, dp, dp, C.POINTER(SolveOpts), C.POINTER(SolveStats)]
    L.beep_hip_solve_device.argtypes = [vp, vp, vp, C.POINTER(SolveOpts), C.POINTER(SolveStats)]
    L.beep_hip_rhs_device.argtypes = [vp, C.c_int, dp, C.c_double, vp]
    L.beep_hip_selftest_mfma.argtypes = [vp, C.c_int, dp, dp, dp]
    L.beep_hip_fmm_symmetry.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, ip, ip]
    L.beep_hip_device_alloc.argtypes = [vp, C.c_size_t, C.POINTER(vp)]
    L.beep_hip_device_free.argtypes = [vp, vp]
    L.beep_hip_memcpy.argtypes = [vp, vp, vp, C.c_size_t, C.c_int]
    L.beep_hip_memory_held.argtypes = [vp, C.POINTER(C.c_size_t)]
    L.beep_hip_synchronize.argtypes = [vp]
    L.beep_hip_gather_rows64.argtypes = [vp, vp, vp, C.c_size_t, vp]
    L.beep_hip_memory_info.argtypes = [vp, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    L.beep_hip_device_memory.argtypes = [C.c_int, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]


def check_host(rc):
    if rc != 0:
        raise BeepError(lib().beep_host_last_error().decode())


def check_hip(ctx, rc):
    if rc != 0:
        raise BeepError(lib().beep_hip_last_error(ctx).decode())


class PoseOpts(C.Structure):
    _fields_ = [
        ("device", C.c_int), ("slots", C.c_int), ("solve_mode", C.c_int), ("tol", C.c_double),
        ("max_iter", C.c_int), ("max_refine", C.c_int), ("inner_tol", C.c_double),
        ("far_mode", C.c_int), ("fmm_order", C.c_int), ("fmm_leaf", C.c_int),
        ("pair_cutoff", C.c_double), ("gram_schmidt", C.c_int), ("keep_fh", C.c_int),
    ]


class PoseResult(C.Structure):
    _fields_ = [
        ("status", C.c_int), ("energy", C.c_double), ("iterations", C.c_int), ("matvecs", C.c_int),
        ("refinements", C.c_int), ("converged", C.c_int), ("rel_residual", C.c_double),
        ("n_pairs", C.c_longlong), ("host_ms", C.c_double), ("device_ms", C.c_double),
        ("slot", C.c_int), ("message", C.c_char * 128),
    ]


def _declare_poses(L):
    L.beep_pose_default_opts.argtypes = [C.POINTER(PoseOpts)]
    L.beep_pose_default_opts.restype = None
    L.beep_poses_create.argtypes = [C.POINTER(PoseOpts), vp, C.c_double, C.c_double, C.c_int, C.c_int,
                                    C.c_int, C.c_double, C.POINTER(vp)]
    L.beep_poses_solve.argtypes = [vp, C.c_int, dp, dp, C.POINTER(PoseResult)]
    L.beep_poses_take_fh.argtypes = [vp, C.c_int, dp, dp]
    L.beep_poses_info.argtypes = [vp, ip, C.POINTER(C.c_longlong), dp, dp]
    L.beep_poses_destroy.argtypes = [vp]
    L.beep_poses_destroy.restype = None
