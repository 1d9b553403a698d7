"""Benchmark: pressure solves/s to 1e-8 relative residual at 256^3 on one MI355X.

Workload: the variable-density pressure-Poisson system on a 256^3 unit cube,
16 seeded bubbles (radius 0.03-0.12, density 1 inside, 1000 outside), Neumann
walls, right-hand side = divergence of a seeded uniform(-1,1) face-velocity
field shifted to zero mean, x0 = 0.  A step is one complete solve to
||b - A x||_2 <= 1e-8 ||b||_2 by CG preconditioned with the additive-correction
V-cycle (2+2 damped Jacobi, coarsest 8^3 exact FP64): FP32 inner solve inside
FP64 iterative refinement (the value), and the plain FP64 solve beside it.
Matrix and right-hand side are resident in HBM before the timed region.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256] [--no-cpu]
"""
import argparse
import json
import os
import re
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "psi-boil_amd"))

import psb  # noqa: E402

HBM_PEAK_GBS = 8000.0
CASE = dict(nsph=16, rmin=0.03, rmax=0.12, rho_in=1.0, rho_out=1000.0, seed_sph=2, seed_vel=1002)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def timed_solves(s, dx, db, steps, warmup, **stop):
    res = None
    for _ in range(warmup):
        dx.zero()
        res = s.solve_device(dx.addr, db.addr, **stop)
    psb.device_synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        dx.zero()
        res = s.solve_device(dx.addr, db.addr, **stop)
    psb.device_synchronize()
    t1 = time.perf_counter()
    return (t1 - t0) / steps, res


def measured_traffic(n):
    """HBM bytes per launch of the FP32 Jacobi sweep at 256^3 from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (calibrated against a
    plain copy), or None."""
    path = os.path.join(ROOT, "profiles", "jacobi_spmv_256_hbm_traffic.md")
    if n != 256 or not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_launch k_jacobi<float,4> (\d+)", open(path).read())
    return int(m.group(1)) if m else None


def cpu_baseline(n, iters_full, sample_iters):
    """Reference CG with diagonal preconditioning (cg.cpp, built from the
    reference sources into oracle/_ref) on one host core: a bounded sample of
    `sample_iters` iterations on the same 256^3 system, scaled to a full solve
    with the iteration count the same algorithm needs on this system."""
    drv = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
    if not os.path.exists(drv):
        return None
    args = [drv, "bench-cg", str(n), str(CASE["nsph"]), repr(CASE["rmin"]), repr(CASE["rmax"]),
            repr(CASE["rho_in"]), repr(CASE["rho_out"]), str(CASE["seed_sph"]),
            str(CASE["seed_vel"]), str(sample_iters)]
    out = subprocess.run(args, capture_output=True, text=True, timeout=600).stdout
    m = re.search(r"solve_s=([0-9.eE+-]+)", out)
    if not m:
        return None
    s_iter = float(m.group(1)) / sample_iters
    return {"value": 1.0 / (s_iter * iters_full), "unit": "solves/s", "cores": 1,
            "kind": "reference",
            "sample": f"reference CG-di (Src/Solver/Linear/Krylov/cg.cpp) {sample_iters} iterations "
                      f"at {n}^3, {s_iter:.3f} s/iteration on 1 core, scaled by the "
                      f"{iters_full} iterations CG-di needs on this system",
            "s_per_iteration": s_iter, "iterations_full": iters_full}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=20)
    ap.add_argument("--profile-kernels", action="store_true",
                    help="only launch the roofline kernels (FP32 Jacobi sweep, FP32/FP64 SpMV) "
                         "and a plain copy, for hardware-counter runs")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported in this build "
              "(the solver has no multi-GPU path: it would shard by 3-D blocks with a "
              "one-cell halo exchange and all-reduced scalars)", file=sys.stderr)
        sys.exit(2)
    n = a.n
    log(f"[bench] generating {n}^3 case")
    d = psb.generate(n, n, n, **CASE)
    s = psb.Solver(n, n, n)
    s.set_matrix(d)
    s.set_geometry(d["dxc"], d["dyc"], d["dzc"])
    if a.profile_kernels:
        for kern, name in ((2, "jacobi_fp32"), (1, "spmv_fp32"), (0, "spmv_fp64"), (3, "copy_fp32")):
            ms, by = s.time_kernel(kern, reps=10)
            log(f"[bench] {name}: {ms * 1e3:.1f} us/launch, {by / 1e6:.1f} MB algorithmic, "
                f"{by / ms / 1e6:.0f} GB/s")
        return
    db = psb.DeviceArray.from_host(d["rhs"])
    dx = psb.DeviceArray(n ** 3)
    common = dict(precond=psb.PSB_PREC_AC, norm=psb.PSB_NORM_L2REL, res_rat=1e-8, res_tol=0.0,
                  max_iter=1000)
    log("[bench] FP64 AC-PCG")
    t64, r64 = timed_solves(s, dx, db, a.steps, a.warmup, precision=psb.PSB_FP64, **common)
    log(f"[bench] fp64 {t64 * 1e3:.2f} ms/solve {r64}")
    log("[bench] mixed AC-PCG, FP32 iterations with FP64 reliable updates")
    tmx, rmx = timed_solves(s, dx, db, a.steps, a.warmup, precision=psb.PSB_MIXED,
                            inner_rtol=1e-2, inner_adapt=2, **common)
    log(f"[bench] mixed {tmx * 1e3:.2f} ms/solve {rmx}")
    x_mixed = dx.download()
    # FP64 reference solution to 1e-12 for the error figure (untimed)
    dx.zero()
    r12 = s.solve_device(dx.addr, db.addr, precision=psb.PSB_FP64, **dict(common, res_rat=1e-12))
    x_ref = dx.download()
    xm, xr = x_mixed - x_mixed.mean(), x_ref - x_ref.mean()
    err = float(np.linalg.norm(xm - xr) / np.linalg.norm(xr))
    # dominant kernel: FP32 damped Jacobi sweep on the fine level
    ms_j, by_j = s.time_kernel(2, reps=50)
    ms_s32, by_s32 = s.time_kernel(1, reps=50)
    ms_s64, by_s64 = s.time_kernel(0, reps=50)
    gbs = lambda by, ms: by / (ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "kernel": "k_jacobi<float,4> (fine-level damped Jacobi sweep)",
            "achieved": round(gbs(by_j, ms_j), 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs(by_j, ms_j) / HBM_PEAK_GBS, 4), "traffic": measured_traffic(n),
            "avg_launch_ms": ms_j, "bytes_per_launch": by_j,
            "spmv_fp32_GBs": round(gbs(by_s32, ms_s32), 1),
            "spmv_fp64_GBs": round(gbs(by_s64, ms_s64), 1)}
    cpu = None
    if not a.no_cpu:
        log("[bench] FP64 CG-di iteration count for the CPU baseline scale")
        dx.zero()
        rdi = s.solve_device(dx.addr, db.addr, precision=psb.PSB_FP64, norm=psb.PSB_NORM_L2REL,
                             res_rat=1e-8, res_tol=0.0, max_iter=100000, graph_chunk=64)
        log(f"[bench] CG-di {rdi}")
        log("[bench] reference CPU sample")
        cpu = cpu_baseline(n, rdi["iterations"], a.cpu_iters)
    out = {
        "metric": "pressure solves/s to 1e-8 rel. residual at 256^3; %HBM roofline; mixed vs FP64",
        "value": round(1.0 / tmx, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(tmx * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 iterations / fp64 residual updates",
        "data": "synthetic (seeded bubbles, seeded face velocities)",
        "config": {"workload": f"pressure Poisson {n}^3, 16 bubbles r 0.03-0.12, density "
                               "ratio 1000, Neumann walls; AC-preconditioned CG (2+2 damped "
                               "Jacobi, 2x coarsening to 8^3, exact FP64 coarsest) to 1e-8; "
                               "FP32 CG with an FP64 true-residual update per 1e-2 drop",
                   "grid": [n, n, n], "parallelism": "single GPU"},
        "fp64_solves_per_s": round(1.0 / t64, 3),
        "fp64_ms_per_solve": round(t64 * 1e3, 3),
        "mixed_over_fp64": round(t64 / tmx, 3),
        "iterations": {"fp64": r64["iterations"], "mixed": rmx["iterations"],
                       "mixed_fp64_updates": rmx["outer"]},
        "rel_residual": {"fp64": r64["rel_l2"], "mixed": rmx["rel_l2"]},
        "error_vs_fp64_1e-12": err,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
