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

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("BEM mat-vecs/s + solve time @163k patches (FP64 vs mixed); solves/s, p99 lat")
HBM_PEAK_GBS = 8000.0
FP32_VECTOR_TFLOPS = 157.3
FP64_VECTOR_TFLOPS = 78.6
FLOP_PER_EVAL = 92  # integrand evaluation (qual point x quad point)


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
    return (t1 - t0) * 1e3 / steps, tm.near_ms_sum / max(tm.count, 1), tm.far_ms_sum / max(tm.count, 1)


def near_bytes(a, coef_bytes):
    """Compulsory bytes of one near-field SpMV launch set: every coefficient
    once, every leaf's column list once, leaf rows + diagonal once, x read
    once and y written once."""
    lp, npt = a["leaf_ptr"], a["nb_ptr"]
    R = np.diff(lp).astype(np.int64)
    C = np.diff(npt).astype(np.int64)
    nnz = int(np.sum(R * C))
    N = int(a["area"].shape[0])
    return nnz * coef_bytes + int(C.sum()) * 4 + N * (4 + 32) + 2 * N * 8 + 2 * N * 8, nnz


def cpu_baseline(b, seconds=15.0):
    """Oracle (oracle/oracle.c, FP64, OpenMP) on a sample of rows of the same
    workload; rate scaled to whole mat-vecs."""
    import oracle
    a = b.scene.arrays()
    o = oracle.Oracle(a, b.kappa, b.Dsolvent)
    N = o.n
    rng = np.random.default_rng(5)
    x = rng.normal(size=2 * N)
    o.diag()
    rows_done, t_used, k = 0, 0.0, 16
    perm = rng.permutation(N)
    while t_used < seconds and rows_done < N:
        rows = perm[rows_done:rows_done + k].astype(np.int32)
        t0 = time.perf_counter()
        o.matvec(x, rows=rows)
        dt = time.perf_counter() - t0
        t_used += dt
        rows_done += len(rows)
        k = min(N - rows_done, max(16, int(len(rows) * min(4.0, (seconds - t_used) / max(dt, 1e-3)))))
        if k <= 0:
            break
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": rows_done / t_used / N, "unit": "mat-vecs/s", "cores": cores, "kind": "port",
            "sample": f"{rows_done} of {N} rows of one FP64 mat-vec (near quadrature + exact "
                      f"1-point far sum, diagonal), {t_used:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="globular")
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--far", default="fmm", choices=["fmm", "direct", "auto"])
    ap.add_argument("--order", type=int, default=5)
    ap.add_argument("--leaf", type=int, default=160)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: every box in this setup has one MI355X, and the "
              "multi-GPU path (octree-subtree sharding, all-gather x / reduce-scatter y per "
              "mat-vec) is a design, not a build.  Run with --gpus 1.", file=sys.stderr)
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
        ms, near_ms, far_ms = time_matvecs(b, d_x, d_y, p, args.warmup, steps)
        rates[p] = {"ms_per_matvec": ms, "near_ms": near_ms, "far_ms": far_ms,
                    "matvecs_per_s": 1e3 / ms, "steps": steps}
        print(f"[bench] {p}: {ms:.3f} ms/matvec (near {near_ms:.3f}, far {far_ms:.3f})",
              file=sys.stderr, flush=True)

    nb64, nnz = near_bytes(a, 32)
    nb32, _ = near_bytes(a, 16)
    head = rates["fp64"]
    roofline = {"bound": "hbm", "achieved": nb64 / (head["near_ms"] * 1e-3) / 1e9,
                "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": nb64 / (head["near_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": None,
                "kernel": "k_near_spmv<double,double>", "bytes_per_launch": nb64}
    extra = {
        "nnz_near_pairs": nnz, "leaves": int(len(lp) - 1), "integrand_evals": n_evals,
        "precompute_ms": pre,
        "integrand_tflops": {p: n_evals * FLOP_PER_EVAL / (pre[p] * 1e-3) / 1e12 for p in pre},
        "integrand_frac_of_vector_peak": {
            "fp64": n_evals * FLOP_PER_EVAL / (pre["fp64"] * 1e-3) / 1e12 / FP64_VECTOR_TFLOPS,
            "fp32": n_evals * FLOP_PER_EVAL / (pre["fp32"] * 1e-3) / 1e12 / FP32_VECTOR_TFLOPS},
        "matvec": rates,
        "near_spmv_fp32_GBs": nb32 / (rates["fp32"]["near_ms"] * 1e-3) / 1e9,
        "host_setup_s": setup_s,
    }

    if not args.no_solve:
        solves = {}
        for mode in ("fp64", "mixed"):
            b.solve(1e-6, 100, mode=mode)
            E = b.calculate_energies()
            ri = dict(b.run_info)
            solves[mode] = {"solve_ms": ri["rhs_ms"] + ri["gmres_ms"], "iterations": ri["iterations"],
                            "matvecs": ri["matvecs"], "refinements": ri["refinements"],
                            "rel_residual": ri["rel_residual"], "energy_kJ_mol": E}
            print(f"[bench] solve {mode}: {solves[mode]}", file=sys.stderr, flush=True)
        solves["mixed_speedup_vs_fp64"] = solves["fp64"]["solve_ms"] / solves["mixed"]["solve_ms"]
        # reference energy: FP64 coefficients, exact direct far field (the
        # explicit-pair + 1-point operator the oracle defines), FP64 GMRES
        if args.far != "direct":
            dev.set_far_field("direct")
            b.solve(1e-8, 100, mode="fp64")
            E_ref = b.calculate_energies()
            dev.set_far_field(args.far, args.order, args.leaf)
        else:
            E_ref = solves["fp64"]["energy_kJ_mol"]
        solves["reference_energy_kJ_mol"] = E_ref
        for mode in ("fp64", "mixed"):
            solves[mode]["energy_rel_err_vs_reference"] = abs(
                solves[mode]["energy_kJ_mol"] - E_ref) / abs(E_ref)
        extra["solve"] = solves

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(b, args.cpu_seconds)
        print(f"[bench] cpu baseline {cpu}", file=sys.stderr, flush=True)

    dev.free(d_x)
    dev.free(d_y)
    out = {
        "metric": METRIC, "value": head["matvecs_per_s"], "unit": "mat-vecs/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": head["ms_per_matvec"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded mesh, charges and x)",
        "config": {"workload": workload, "patches": N, "quad_rule": 6, "qual_rule": 0,
                   "nbsize": b.nbsize, "kappa": b.kappa, "eps_in": 2.0, "eps_out": 80.0,
                   "far_field": (f"Chebyshev FMM order {args.order}, leaf target {args.leaf}"
                                 if args.far == "fmm" else args.far)},
        "roofline": roofline, "cpu_baseline": cpu, "extra": extra,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
