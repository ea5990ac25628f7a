#!/usr/bin/env python3
"""Benchmark of the BEM mat-vec inside the GMRES solve on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload globular]

A step is one BEM mat-vec y = K x (near-field block SpMV + diagonal + far
field) over device-resident x, y at the named workload (default: the
163,842-patch globular surface with 2,000 charges).  The headline value is the
FP64 mat-vec rate; FP32 (mixed-solve inner) and FP32-coefficient/FP64-sum
(refinement residual) rates, precompute and solve times (plain FP64 vs mixed)
are reported beside it.  Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

# hardware queues of the HIP runtime (read when it starts, before any HIP
# call): the solve server runs two streams per slot, and with HIP's default
# of 4 queues a small request waits behind other slots' long kernels
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < 16:
    os.environ["GPU_MAX_HW_QUEUES"] = "16"

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("BEM mat-vecs/s + solve time @163k patches (FP64 vs mixed); solves/s, p99 lat")
HBM_PEAK_GBS = 8000.0
FP32_VECTOR_TFLOPS = 157.3
FP64_VECTOR_TFLOPS = 78.6
FLOP_PER_EVAL = 92  # integrand evaluation (qual point x quad point)
MFMA_FP64_TFLOPS = 78.6   # dense matrix peaks of one MI355X
MFMA_FP32_TFLOPS = 157.3


TUNE_LEAVES = (24, 32, 40, 64, 96, 160)


def build_workload(name, **kw):
    from beep3_amd import configs
    if name == "globular":
        return configs.globular(**kw), "globular surface r=25 A, 163,842 patches, 2,000 charges"
    if name == "off_centre":
        return configs.off_centre(**kw), "icosphere r=20 A, 40,962 patches, 200 charges"
    if name == "born":
        return configs.born_ion(**kw), "icosphere r=2 A, 2,562 patches, 1 charge"
    raise SystemExit(f"unknown workload {name}")


def time_matvecs(b, d_x, d_y, precision, warmup, steps):
    dev = b.dev
    for _ in range(warmup):
        dev.matvec_device(d_x, d_y, precision)
    dev.synchronize()
    dev.timing_reset()
    t0 = time.perf_counter()
    for _ in range(steps):
        dev.matvec_device(d_x, d_y, precision)
    dev.synchronize()
    t1 = time.perf_counter()
    tm = dev.timing()
    k = max(tm.count, 1)
    time_matvecs.m2l_ms = tm.m2l_ms_sum / k
    return ((t1 - t0) * 1e3 / steps, tm.near_ms_sum / k, tm.far_ms_sum / k, tm.total_ms_sum / k)


def near_bytes(a, coef_bytes):
    """Bytes of one near-field SpMV launch set.

    algorithmic: the per-pair figure of the reference layout (4 coefficients
    and one 4-byte column index per pair, local_integrations.h:17-25) plus x
    read once (2N values of 8 B) and y written once (2N x 8 B);
    compulsory: what this block layout must move (every coefficient once,
    one column index per (leaf, column) instead of per pair, leaf rows and
    the diagonal once, x once, y once)."""
    lp, npt = a["leaf_ptr"], a["nb_ptr"]
    R = np.diff(lp).astype(np.int64)
    C = np.diff(npt).astype(np.int64)
    nnz = int(np.sum(R * C))
    N = int(a["area"].shape[0])
    algorithmic = nnz * (coef_bytes + 4) + 2 * N * 8 + 2 * N * 8
    compulsory = nnz * coef_bytes + int(C.sum()) * 4 + N * (4 + 32) + 2 * N * 8 + 2 * N * 8
    return algorithmic, compulsory, nnz


# HBM bytes per launch on the 163,842-patch globular surface, from rocprofv3
# --pmc FETCH_SIZE and WRITE_SIZE (profiles/matvec_globular_163842_fmm5_pmc.md):
# FETCH_SIZE x 2.0 and WRITE_SIZE x 1.0, the factors that make the counters
# read 1 GiB for a 1 GiB device copy.  "spmv": the FP64 near-SpMV launch set
# (k_near_spmv_part<double,double,*> + k_near_spmv_sum<double>; the depth
# variants read the same); "fmm": one whole FP64 mat-vec with the Chebyshev
# far field (SpMV, far field and the residual leaf pass).
MEASURED_TRAFFIC = {("globular", "spmv", "fp64"): 8.309e9,
                    ("globular", "fmm", "fp64"): 9.390e9}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(b, seconds=15.0, solve_matvecs=None):
    """The reference's CPU path restated in C++/OpenMP (oracle/cpu_ref.cpp)
    on the host's cores: FP32 24-byte LocalIntegrations applied by scatter
    with FP64 sums (beep.cpp:768-956, local_integrations.h:123-130) and the
    FP64 plane-wave FMM of order 9 with the 9-node quadrature
    (fmm_bem_hybrid.h:184-642, the reference's <9,9,67>).  The value is whole
    mat-vecs per second (every record is walked).  The precompute is sampled
    (every 16th near-list leaf) and scaled to all records; the solve time is
    the tree + precompute + the mat-vecs the GPU's FP64 solve needed."""
    import oracle
    a = b.host_scene().arrays()
    o = oracle.Oracle(a, b.kappa, b.Dsolvent)
    N = o.n
    stride = 16
    t0 = time.perf_counter()
    c = oracle.CpuReference(o, 9, 9, 40, stride=stride)
    create_s = time.perf_counter() - t0
    info = c.info
    x = np.random.default_rng(5).normal(size=2 * N)
    times = []
    t_used = 0.0
    while len(times) < 3 or (t_used < seconds and len(times) < 20):
        _, ms = c.matvec(x)
        times.append(ms)
        t_used += ms * 1e-3
    c.close()
    ms = float(np.median(times))
    scale = info["records"] / max(info["records_computed"], 1)
    precompute_ms = info["near_ms"] * scale
    setup_ms = info["setup_ms"] - info["near_ms"] + precompute_ms
    out = {"value": 1e3 / ms, "unit": "mat-vecs/s", "cores": int(info["threads"]),
           "kind": "port", "cpu_model": cpu_model(), "nproc": os.cpu_count(),
           "sample": (f"{len(times)} whole FP64 mat-vecs at {N} patches (median {ms:.1f} ms); "
                      f"precompute of {info['records_computed']} of {info['records']} 24-byte "
                      f"records ({info['near_ms']:.0f} ms, x{scale:.1f} to all records); "
                      f"{create_s + t_used:.1f} s of CPU work"),
           "matvec_ms": ms, "precompute_ms_scaled": precompute_ms,
           "setup_ms_scaled": setup_ms, "records": info["records"],
           "residual_pairs": info["residual_pairs"], "fmm": "plane-wave order 9, 9 nodes, leaf 40"}
    if solve_matvecs:
        out["solve_ms_estimated"] = setup_ms + solve_matvecs * ms
        out["solve_matvecs"] = solve_matvecs
    return out


def serving_bench(seconds, slots, seq_requests, rate=50.0):
    """Poisson stream of single-molecule solves (rate per second over
    `seconds`) through the continuous-batching server, then the first
    `seq_requests` of the same stream as sequential plain FP64 solves (one
    slot, one at a time)."""
    from beep3_amd import configs
    from beep3_amd.serve import Server

    sched = configs.serving_schedule(rate=rate, duration=seconds, seed=11)
    t0 = time.perf_counter()
    meshes = configs.serving_meshes()
    gen_s = time.perf_counter() - t0
    print(f"[bench] serving: {len(sched)} requests over {seconds:.0f}s, mesh pool {gen_s:.1f}s",
          file=sys.stderr, flush=True)

    def run(srv, reqs, paced):
        ids, res = {}, []
        t_start = srv.clock()
        for q in reqs:
            if paced:
                wait = t_start + q["t"] - srv.clock()
                if wait > 0:
                    time.sleep(wait)
            sc, _ = configs.serving_request(meshes, q)
            ids[srv.submit(sc)] = q
            if not paced:  # sequential: one request at a time
                r = None
                while r is None:
                    r = srv.poll(timeout=5.0)
                res.append(r)
            else:
                r = srv.poll(timeout=0.0)
                while r is not None:
                    res.append(r)
                    r = srv.poll(timeout=0.0)
        while len(res) < len(reqs):
            r = srv.poll(timeout=5.0)
            if r is not None:
                res.append(r)
        t_end = max(r["t_done"] for r in res)
        return res, t_end - t_start

    srv = Server(slots=slots, mode="mixed", far_field="auto", fmm_order=5,
                 fmm_leaf=64)
    res, span = run(srv, sched, paced=True)
    srv.close()
    ok = [r for r in res if r["status"] == 0]
    lat = np.array([r["latency_s"] for r in ok]) * 1e3
    by_level = {}
    for r in ok:
        by_level.setdefault(r["n_patches"], []).append(r["latency_s"] * 1e3)
    out = {"requests": len(sched), "solved": len(ok), "failed": len(res) - len(ok),
           "failures": sorted({r["message"] for r in res if r["status"] != 0}),
           "slots": slots, "arrival_rate_per_s": rate, "span_s": span,
           "solves_per_s": len(ok) / span, "latency_ms_p50": float(np.percentile(lat, 50)),
           "latency_ms_p99": float(np.percentile(lat, 99)),
           "latency_ms_p50_by_patches": {str(k): float(np.percentile(v, 50)) for k, v in sorted(by_level.items())},
           "phase_ms_median_by_patches": {
               str(k): {f: float(np.median([r[f] for r in ok if r["n_patches"] == k]))
                        for f in ("host_ms", "upload_ms", "precompute_ms", "solve_ms")}
               for k in sorted(by_level)},
           "host_ms_mean": float(np.mean([r["host_ms"] for r in ok])),
           "device_ms_mean": float(np.mean([r["device_ms"] for r in ok])),
           "max_device_bytes_per_slot": float(max(r["bytes"] for r in ok)),
           "unconverged": sum(1 for r in ok if not r["converged"])}
    print(f"[bench] serving: {out}", file=sys.stderr, flush=True)
    # capacity: the same requests submitted at once (closed loop)
    srv = Server(slots=slots, mode="mixed", far_field="auto", fmm_order=5,
                 fmm_leaf=64)
    cres, cspan = run(srv, [dict(q, t=0.0) for q in sched], paced=True)
    srv.close()
    out["capacity_solves_per_s"] = sum(1 for r in cres if r["status"] == 0) / cspan
    print(f"[bench] serving capacity: {out['capacity_solves_per_s']:.1f} solves/s",
          file=sys.stderr, flush=True)
    srv = Server(slots=1, mode="fp64", far_field="auto", fmm_order=5, fmm_leaf=64)
    sres, sspan = run(srv, sched[:seq_requests], paced=False)
    srv.close()
    sok = [r for r in sres if r["status"] == 0]
    service = np.array([(r["t_done"] - r["t_submit"]) * 1e3 for r in sok])
    out["sequential_plain_fp64"] = {
        "requests": len(sres), "solves_per_s": len(sok) / sspan,
        "service_ms_p50": float(np.percentile(service, 50)),
        "service_ms_p99": float(np.percentile(service, 99))}
    print(f"[bench] sequential plain: {out['sequential_plain_fp64']}", file=sys.stderr, flush=True)
    return out


def pose_bench(n_poses, slots, seq_poses):
    """Two instances of the 40,962-patch surface (a_lm seed 7) in n_poses
    seeded rigid poses (FP32 mat-vecs inside FP64 refinement):
    * the pose batch: the mesh's own near field computed once and shared by
      both instances of every pose, inter-instance pairs per pose;
    * the same poses through the server, each with the reference's split
      (one octree over the pose, every explicit pair integrated per pose);
    then seq_poses of them as sequential plain FP64 server solves."""
    from beep3_amd import configs
    from beep3_amd.poses import PoseBatch
    from beep3_amd.serve import Server

    mesh = configs.pose_mesh()
    rng = np.random.default_rng(7)  # the generator of configs.pose_scenes
    poses = [configs.random_pose(rng) for _ in range(n_poses)]
    t0 = time.perf_counter()
    pb = PoseBatch(mesh, slots=slots, mode="mixed", far_field="fmm", fmm_order=5, fmm_leaf=64)
    setup_s = time.perf_counter() - t0
    info = pb.info()
    t0 = time.perf_counter()
    bres = pb.solve(poses)
    bspan = time.perf_counter() - t0
    pb.close()
    bok = [r for r in bres if r["status"] == 0]
    batch = {"poses": n_poses, "solved": len(bok), "slots": slots, "span_s": bspan,
             "poses_per_s": len(bok) / bspan, "setup_s_once": setup_s,
             "poses_per_s_with_setup": len(bok) / (bspan + setup_s),
             "mesh_patches": info["n_patches"], "intra_pairs_shared": info["intra_pairs"],
             "pair_cutoff_A": info["cutoff"],
             "inter_pairs_median": float(np.median([r["n_pairs"] for r in bok])),
             "iterations_mean": float(np.mean([r["iterations"] for r in bok])),
             "phase_ms_median": {f: float(np.median([r[f] for r in bok])) for f in ("host_ms", "device_ms")},
             "unconverged": sum(1 for r in bok if not r["converged"]),
             "failures": sorted({r["message"] for r in bres if r["status"] != 0})}
    print(f"[bench] pose batch: {batch}", file=sys.stderr, flush=True)
    scenes = configs.pose_scenes(mesh, n_poses)

    def run(srv, scs, one_at_a_time):
        res = []
        t0 = srv.clock()
        for sc in scs:
            srv.submit(sc)
            if one_at_a_time:
                r = None
                while r is None:
                    r = srv.poll(timeout=5.0)
                res.append(r)
        while len(res) < len(scs):
            r = srv.poll(timeout=5.0)
            if r is not None:
                res.append(r)
        return res, max(r["t_done"] for r in res) - t0

    srv = Server(slots=slots, mode="mixed", far_field="fmm", fmm_order=5, fmm_leaf=64)
    res, span = run(srv, scenes, False)
    srv.close()
    ok = [r for r in res if r["status"] == 0]
    # the two splits, pose by pose (the server returns in completion order;
    # request ids follow submission order from 1)
    by_id = {r["id"]: r for r in res}
    dE = [abs(b["energy"] - by_id[k + 1]["energy"]) / abs(by_id[k + 1]["energy"])
          for k, b in enumerate(bres) if b["status"] == 0 and by_id.get(k + 1, {}).get("status") == 0]
    batch["energy_rel_diff_vs_reference_split"] = {"median": float(np.median(dE)), "max": float(np.max(dE))}
    lat = np.array([r["latency_s"] for r in ok]) * 1e3
    E = np.array([r["energy"] for r in ok])
    out = {"pose_batch": batch,
           "poses": n_poses, "solved": len(ok), "patches_per_pose": ok[0]["n_patches"] if ok else 0,
           "slots": slots, "span_s": span, "poses_per_s": len(ok) / span,
           "latency_ms_p50": float(np.percentile(lat, 50)),
           "latency_ms_p99": float(np.percentile(lat, 99)),
           "iterations_mean": float(np.mean([r["iterations"] for r in ok])),
           "phase_ms_median": {f: float(np.median([r[f] for r in ok]))
                               for f in ("host_ms", "upload_ms", "precompute_ms", "solve_ms")},
           "unconverged": sum(1 for r in ok if not r["converged"]),
           "energy_kJ_mol_range": [float(E.min()), float(E.max())],
           "failures": sorted({r["message"] for r in res if r["status"] != 0})}
    print(f"[bench] poses through the server (reference split): "
          f"{ {k: v for k, v in out.items() if k != 'pose_batch'} }", file=sys.stderr, flush=True)
    seq = configs.pose_scenes(mesh, seq_poses)
    srv = Server(slots=1, mode="fp64", far_field="fmm", fmm_order=5, fmm_leaf=64)
    sres, sspan = run(srv, seq, True)
    srv.close()
    out["sequential_plain_fp64"] = {"poses": seq_poses, "poses_per_s": len(sres) / sspan}
    print(f"[bench] poses sequential plain: {out['sequential_plain_fp64']}", file=sys.stderr,
          flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="globular")
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--no-fly", action="store_true",
                    help="skip the on-the-fly near-field timing")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--far", default="fmm", choices=["fmm", "pw", "direct", "auto"])
    ap.add_argument("--order", type=int, default=5)
    ap.add_argument("--leaf", type=int, default=40)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=6.0)
    ap.add_argument("--serve-slots", type=int, default=8)
    ap.add_argument("--serve-seq", type=int, default=40)
    ap.add_argument("--poses", type=int, default=256)
    ap.add_argument("--copy-gib", type=float, default=0.0,
                    help="one device-to-device copy of this many GiB first (a plain copy of "
                         "known size, to calibrate hardware byte counters)")
    ap.add_argument("--gather-gib", type=float, default=0.0,
                    help="one random gather of this many GiB of 64-byte rows first (every "
                         "row read once in random order, plus 4-byte indices), to check the "
                         "counter factors on scattered reads")
    ap.add_argument("--pose-seq", type=int, default=16)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the multi-GPU path (octree-subtree sharding, "
              "one all-gather of x per mat-vec, each GPU writing its own rows of y) is a "
              "design, not a build.  Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    t_setup = time.perf_counter()
    b, workload = build_workload(args.workload, far_field=args.far, fmm_order=args.order,
                                 fmm_leaf=args.leaf)
    b.init(precisions=["fp64", "fp32"])
    setup_s = time.perf_counter() - t_setup
    print(f"[bench] setup done in {setup_s:.1f}s: {b.run_info}", file=sys.stderr, flush=True)
    a = b.scene.arrays()
    N = b.n
    dev = b.dev
    if args.copy_gib > 0:
        nbytes = int(args.copy_gib * 2**30)
        src, dst = dev.alloc(nbytes), dev.alloc(nbytes)
        dev.copy_device(dst, src, nbytes)
        dev.synchronize()
        dev.free(src)
        dev.free(dst)
    if args.gather_gib > 0:
        rows = int(args.gather_gib * 2**30) // 64
        perm = np.random.default_rng(1).permutation(rows).astype(np.int32)
        g_src, g_dst, g_idx = dev.alloc(rows * 64), dev.alloc(rows * 64), dev.alloc(rows * 4)
        dev.upload(g_idx, perm)
        dev.gather_rows64(g_src, g_idx, rows, g_dst)
        dev.synchronize()
        print(f"[bench] gather: {rows} rows of 64 B, reads {rows * 64 + rows * 4} B, writes {rows * 64} B",
              file=sys.stderr, flush=True)
        for ptr in (g_src, g_dst, g_idx):
            dev.free(ptr)
    # what a plain device-to-device copy sustains (read + write bytes over its
    # time, best of 5 copies of 2 GiB): the practical HBM ceiling beside the
    # 8 TB/s of the data sheet
    copy_bytes = 2 << 30
    c_src, c_dst = dev.alloc(copy_bytes), dev.alloc(copy_bytes)
    dev.copy_device(c_dst, c_src, copy_bytes)
    dev.synchronize()
    copy_s = []
    for _ in range(5):
        t0 = time.perf_counter()
        dev.copy_device(c_dst, c_src, copy_bytes)
        dev.synchronize()
        copy_s.append(time.perf_counter() - t0)
    dev.free(c_src)
    dev.free(c_dst)
    copy_GBs = 2 * copy_bytes / min(copy_s) / 1e9
    tune = None
    if args.far in ("fmm", "pw") and not args.no_tune:
        # leaf-size autotuner on the FP64 mat-vec at this order
        best, table = dev.autotune_leaf("fp64", TUNE_LEAVES, reps=3)
        tune = {"leaf_ms_fp64": table, "chosen_leaf": best}
        args.leaf = best
        # near-SpMV load depth at the default cap, then the residency cap at
        # that depth (room for the concurrent far field), per precision
        for p in ("fp64", "fp32", "fp32_resid"):
            depth, dtable = dev.autotune_spmv_depth(p, (1, 2, 4), reps=3)
            cap, ctable = dev.autotune_spmv(p, (0, 1, 2, 3, 4, 6), reps=3)
            tune[f"spmv_{p}"] = {"depth_ms": dtable, "depth": depth,
                                 "blocks_per_cu_ms": ctable, "blocks_per_cu": cap}
        print(f"[bench] autotune: {tune}", file=sys.stderr, flush=True)

    # precompute timings (integrand throughput)
    pre = {}
    for p in ("fp64", "fp32"):
        dev.synchronize()
        t0 = time.perf_counter()
        dev.precompute(b.kappa, p)
        dev.synchronize()
        pre[p] = (time.perf_counter() - t0) * 1e3
    qn = np.diff(a["qual_ptr"]).astype(np.int64)
    dn = np.diff(a["quad_ptr"]).astype(np.int64)
    lp, li, npt, ni = a["leaf_ptr"], a["leaf_idx"], a["nb_ptr"], a["nb_idx"]
    n_evals = 0
    for L in range(len(lp) - 1):
        n_evals += int(qn[li[lp[L]:lp[L + 1]]].sum()) * int(dn[ni[npt[L]:npt[L + 1]]].sum())
    print(f"[bench] precompute ms {pre}, integrand evals {n_evals:.3e}", file=sys.stderr, flush=True)

    rng = np.random.default_rng(1)
    x = rng.normal(size=2 * N)
    d_x = dev.alloc(16 * N)
    d_y = dev.alloc(16 * N)
    dev.upload(d_x, x)

    rates = {}
    for p, steps in (("fp64", args.steps), ("fp32", max(3, args.steps // 2)),
                     ("fp32_resid", max(3, args.steps // 2))):
        ms, near_ms, far_ms, tot_ms = time_matvecs(b, d_x, d_y, p, args.warmup, steps)
        rates[p] = {"ms_per_matvec": ms, "near_ms": near_ms, "far_ms": far_ms,
                    "device_ms": tot_ms, "m2l_ms": time_matvecs.m2l_ms,
                    "matvecs_per_s": 1e3 / ms, "steps": steps}
        print(f"[bench] {p}: {ms:.3f} ms/matvec (device {tot_ms:.3f}: near {near_ms:.3f} "
              f"concurrent with far {far_ms:.3f})", file=sys.stderr, flush=True)
    # the ABI's host-vector mat-vec (beep_hip_matvec): x up and y down over
    # PCIe around the same device work (not the headline value)
    host_ms = {}
    for p in ("fp64", "fp32"):
        b.dev.matvec(x, p)
        t0 = time.perf_counter()
        for _ in range(5):
            b.dev.matvec(x, p)
        host_ms[p] = (time.perf_counter() - t0) / 5 * 1e3
    # the near SpMV alone (far field serialised after it, grid uncapped: the
    # SpMV has the whole chip and memory system), for its roofline
    # at each load depth; the fastest depth gives the roofline figure
    dev.set_overlap(False)
    dev.set_spmv_residency(0)
    alone, m2l_alone = {}, {}
    for dd in (1, 2):
        dev.set_spmv_depth(dd)
        _, n64, f64, _ = time_matvecs(b, d_x, d_y, "fp64", 1, max(3, args.steps // 2))
        m2l_alone["fp64"] = time_matvecs.m2l_ms
        _, n32, f32, _ = time_matvecs(b, d_x, d_y, "fp32", 1, max(3, args.steps // 2))
        m2l_alone["fp32"] = time_matvecs.m2l_ms
        alone[dd] = (n64, f64, n32, f32)
    best_d64 = min(alone, key=lambda k: alone[k][0])
    best_d32 = min(alone, key=lambda k: alone[k][2])
    near_alone, far_alone = alone[best_d64][0], alone[best_d64][1]
    near32_alone, far32_alone = alone[best_d32][2], alone[best_d32][3]
    # back to the tuned settings of every precision (or the library defaults);
    # a tuner run over a single candidate sets one precision's value
    dev.set_spmv_residency(1)
    dev.set_spmv_depth(1)
    for p in (("fp64", "fp32", "fp32_resid") if tune else ()):
        dev.autotune_spmv_depth(p, (tune[f"spmv_{p}"]["depth"],), reps=1)
        dev.autotune_spmv(p, (tune[f"spmv_{p}"]["blocks_per_cu"],), reps=1)
    dev.set_overlap(True)

    nb64, cb64, nnz = near_bytes(a, 32)
    nb32, cb32, _ = near_bytes(a, 16)
    head = rates["fp64"]
    fi = dev.far_info()
    # the whole FP64 mat-vec as the timed step runs it: the near SpMV at the
    # tuned depth and residency cap beside the far field on the second
    # stream; "mat-vec achieved" = (near + far algorithmic bytes) over the
    # HIP-event time of one mat-vec in the timed loop
    d64 = tune["spmv_fp64"]["depth"] if tune else 1
    cap64 = tune["spmv_fp64"]["blocks_per_cu"] if tune else 1
    far64 = fi["bytes_fp64"]
    t_mv = head["device_ms"]
    roofline = {"bound": "hbm", "achieved": (nb64 + far64) / (t_mv * 1e-3) / 1e9,
                "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": (nb64 + far64) / (t_mv * 1e-3) / 1e9 / HBM_PEAK_GBS,
                "traffic": MEASURED_TRAFFIC.get((args.workload, args.far, "fp64")),
                "kernel": (f"whole FP64 mat-vec as timed: k_near_spmv_part<double,double,{4 * d64}> "
                           f"(cap {cap64} block(s)/CU) + k_near_spmv_sum<double> beside the "
                           f"far field on a second stream"),
                "near_bytes": nb64, "far_bytes": far64, "bytes_per_launch": nb64 + far64,
                "launch_ms": t_mv, "compulsory_bytes": cb64 + far64}
    near_alone_roofline = {
        "kernel": f"k_near_spmv_part<double,double,{4 * best_d64}> + k_near_spmv_sum<double>, "
                  f"uncapped, far field serialised after it",
        "achieved": nb64 / (near_alone * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": nb64 / (near_alone * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "traffic": MEASURED_TRAFFIC.get((args.workload, "spmv", "fp64")),
        "bytes_per_launch": nb64, "launch_ms": near_alone, "compulsory_bytes": cb64,
        "compulsory_GBs": cb64 / (near_alone * 1e-3) / 1e9}
    m2l = None
    if fi["mode"] == 1 and fi["m2l_pairs"] > 0:
        # real GEMM flops of M2L: 3 products of the P^3 x P^3 table with a
        # source expansion per box pair, over the k_m2l + k_m2l_sum time of a
        # serialised mat-vec (and, beside it, of the overlapped one, where
        # the SpMV shares the chip)
        NP = fi["order"] ** 3
        flops = 2.0 * 3 * NP * NP * fi["m2l_pairs"]
        m2l = {"pairs": fi["m2l_pairs"], "gflop_per_launch": flops / 1e9}
        for p, peak in (("fp64", MFMA_FP64_TFLOPS), ("fp32", MFMA_FP32_TFLOPS)):
            ms = m2l_alone.get(p, 0.0)
            if ms > 0:
                m2l[p] = {"ms": ms, "tflops": flops / (ms * 1e-3) / 1e12,
                          "frac_of_mfma_peak": flops / (ms * 1e-3) / 1e12 / peak,
                          "ms_overlapped": rates[p]["m2l_ms"]}
    extra = {
        "near_spmv_alone_roofline": near_alone_roofline,
        "far_plan": fi,
        "m2l_mfma": m2l,
        "nnz_near_pairs": nnz, "leaves": int(len(lp) - 1), "integrand_evals": n_evals,
        "precompute_ms": pre,
        "integrand_tflops": {p: n_evals * FLOP_PER_EVAL / (pre[p] * 1e-3) / 1e12 for p in pre},
        "integrand_frac_of_vector_peak": {
            "fp64": n_evals * FLOP_PER_EVAL / (pre["fp64"] * 1e-3) / 1e12 / FP64_VECTOR_TFLOPS,
            "fp32": n_evals * FLOP_PER_EVAL / (pre["fp32"] * 1e-3) / 1e12 / FP32_VECTOR_TFLOPS},
        "matvec": rates,
        "serial_phases_ms": {"fp64": {"near": near_alone, "far": far_alone},
                             "fp32": {"near": near32_alone, "far": far32_alone}},
        "near_spmv_alone_ms_by_depth": {dd: {"fp64": v[0], "fp32": v[2]} for dd, v in alone.items()},
        "near_spmv_fp32_GBs": nb32 / (near32_alone * 1e-3) / 1e9,
        "matvec_host_vectors_ms": host_ms,
        "host_setup_s": setup_s,
        "hbm_copy_GBs": copy_GBs,
        # measured HBM bytes of the FP64 SpMV over its time, against the copy
        "near_spmv_fp64_hbm_GBs": (MEASURED_TRAFFIC.get((args.workload, "spmv", "fp64")) or nb64)
                                  / (near_alone * 1e-3) / 1e9,
        "near_spmv_fp64_frac_of_copy": (MEASURED_TRAFFIC.get((args.workload, "spmv", "fp64")) or nb64)
                                       / (near_alone * 1e-3) / 1e9 / copy_GBs,
        "autotune": tune,
    }

    if not args.no_fly:
        # the near field on the fly (no stored coefficients; every mat-vec
        # integrates the explicit pairs): a second context on the same
        # geometry, tree and far field, timed per mat-vec and checked against
        # the stored-coefficient mat-vec on the same x
        from beep3_amd.device import Device

        fd = Device(0)
        try:
            fd.set_far_field(b.far_field, b.fmm_order, args.leaf)
            fd.set_near_mode("on_the_fly")
            sc_ = b.scene
            if sc_.host_points:
                fd.set_geometry(a)
            else:
                fd.set_surface(a, sc_.surface(), b.qual, b.quad)
            fd.set_tree(a)
            fly = {}
            fx, fy = fd.alloc(16 * N), fd.alloc(16 * N)
            fd.upload(fx, x)
            for p in ("fp64", "fp32"):
                fd.precompute(b.kappa, p)
                fd.matvec_device(fx, fy, p)  # plan and tables for this precision
                fd.synchronize()
                t0 = time.perf_counter()
                for _ in range(3):
                    fd.matvec_device(fx, fy, p)
                fd.synchronize()
                ms = (time.perf_counter() - t0) * 1e3 / 3
                y_fly = fd.download(np.empty(2 * N), fy)
                dev.matvec_device(d_x, d_y, p)
                y_ref = dev.download(np.empty(2 * N), d_y)
                fly[p] = {"ms_per_matvec": ms,
                          "integrand_tflops": n_evals * FLOP_PER_EVAL / (ms * 1e-3) / 1e12,
                          "rel_diff_vs_stored": float(np.linalg.norm(y_fly - y_ref) / np.linalg.norm(y_ref)),
                          "device_bytes_held": fd.memory_held()}
            fly["stored_context_bytes_held"] = dev.memory_held()
            extra["near_on_the_fly"] = fly
            print(f"[bench] near field on the fly: {fly}", file=sys.stderr, flush=True)
        finally:
            fd.close()

    if not args.no_solve:
        solves = {}
        # the solve on device vectors (RHS on the device, GMRES from a zero
        # guess): no PCIe transfer of b or x inside the timed region
        charges = b.scene.arrays()["charges"]
        d_rhs, d_sol = dev.alloc(16 * N), dev.alloc(16 * N)
        zero = np.zeros(2 * N)

        def device_solve(mode, **kw):
            dev.upload(d_sol, zero)
            dev.synchronize()
            t0 = time.perf_counter()
            dev.rhs_device(charges, b.Dsolvent, d_rhs)
            t1 = time.perf_counter()
            st = dev.solve_device(d_rhs, d_sol, tol=1e-6, max_iter=100, mode=mode, **kw)
            t2 = time.perf_counter()
            x = dev.download(np.empty(2 * N), d_sol)
            return (t1 - t0) * 1e3, (t2 - t1) * 1e3, st, x

        for mode in ("fp64", "mixed"):  # untimed: GMRES buffers and step graphs
            device_solve(mode)
        fh = {}
        # mixed: inner FP32 GMRES to half the outer tolerance (default), and
        # the looser inner tolerances of a classic refinement
        # mixed_resid64: the refinement residual with the FP64 coefficient set
        # (needs both precomputes; the default residual uses the FP32 set with
        # FP64 sums and far field, as the reference's own FP32 LocalIntegrations)
        for mode, kw in (("fp64", {}), ("mixed", {}), ("mixed_inner1e-3", {"inner_tol": 1e-3}),
                         ("mixed_inner1e-4", {"inner_tol": 1e-4}),
                         ("mixed_resid64", {"resid_fp64": True})):
            rhs_ms, gm_ms, st, x = device_solve(mode.split("_")[0], **kw)
            b.f, b.h = x[:N].copy(), x[N:].copy()
            E = b.calculate_energies()
            fh[mode] = (b.f.copy(), b.h.copy())
            solves[mode] = {"solve_ms": rhs_ms + gm_ms, "rhs_ms": rhs_ms, "gmres_ms": gm_ms,
                            "iterations": st["iterations"], "matvecs": st["matvecs"],
                            "refinements": st["refinements"],
                            "rel_residual": st["rel_residual"], "energy_kJ_mol": E}
            print(f"[bench] solve {mode}: {solves[mode]}", file=sys.stderr, flush=True)
        dev.free(d_rhs)
        dev.free(d_sol)
        solves["mixed_speedup_vs_fp64"] = solves["fp64"]["solve_ms"] / solves["mixed"]["solve_ms"]
        # per-solve setup on the device: singular terms + near coefficients of
        # the precision the solve needs (FP64 for the plain solve, FP32 for mixed)
        for mode, p in (("fp64", "fp64"), ("mixed", "fp32")):
            solves[mode]["with_precompute_ms"] = solves[mode]["solve_ms"] + pre[p]
        solves["mixed_speedup_vs_fp64_with_precompute"] = (
            solves["fp64"]["with_precompute_ms"] / solves["mixed"]["with_precompute_ms"])
        # reference energy: FP64 coefficients, exact direct far field (the
        # explicit-pair + 1-point operator the oracle defines), FP64 GMRES
        if args.far != "direct":
            # the near coefficients follow the far-field plan
            dev.set_far_field("direct")
            dev.precompute(b.kappa, "fp64")
            b.solve(1e-8, 100, mode="fp64")
            E_ref = b.calculate_energies()
            f_ref, h_ref = b.f.copy(), b.h.copy()
            dev.set_far_field(args.far, args.order, args.leaf)
            for p in ("fp64", "fp32"):
                dev.precompute(b.kappa, p)
        else:
            E_ref = solves["fp64"]["energy_kJ_mol"]
            f_ref, h_ref = fh["fp64"]
        solves["reference_energy_kJ_mol"] = E_ref
        for mode in fh:
            solves[mode]["energy_rel_err_vs_reference"] = abs(
                solves[mode]["energy_kJ_mol"] - E_ref) / abs(E_ref)
            solves[mode]["f_rel_l2_err_vs_reference"] = float(
                np.linalg.norm(fh[mode][0] - f_ref) / np.linalg.norm(f_ref))
            solves[mode]["h_rel_l2_err_vs_reference"] = float(
                np.linalg.norm(fh[mode][1] - h_ref) / np.linalg.norm(h_ref))
        extra["solve"] = solves

    dev.free(d_x)
    dev.free(d_y)
    if not args.no_serve:
        b.close()  # free the headline workload's device memory first
        extra["serving"] = serving_bench(args.serve_seconds, args.serve_slots, args.serve_seq)
    if args.poses > 0:
        if args.no_serve:
            b.close()
        extra["poses"] = pose_bench(args.poses, args.serve_slots, args.pose_seq)

    cpu = None
    if not args.no_cpu:
        mv = extra.get("solve", {}).get("fp64", {}).get("matvecs")
        cpu = cpu_baseline(b, args.cpu_seconds, mv)
        print(f"[bench] cpu baseline {cpu}", file=sys.stderr, flush=True)

    out = {
        "metric": METRIC, "value": head["matvecs_per_s"], "unit": "mat-vecs/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": head["ms_per_matvec"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded mesh, charges and x)",
        "config": {"workload": workload, "patches": N, "quad_rule": 6, "qual_rule": 0,
                   "nbsize": b.nbsize, "kappa": b.kappa, "eps_in": 2.0, "eps_out": 80.0,
                   "far_field": (f"Chebyshev FMM order {args.order}, leaf target {args.leaf}"
                                 if args.far == "fmm" else
                                 f"plane-wave FMM order {args.order}, leaf target {args.leaf}"
                                 if args.far == "pw" else args.far)},
        "roofline": roofline, "cpu_baseline": cpu, "extra": extra,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
