"""Benchmark: pressure solves/s to 1e-8 relative residual at 256^3 on one MI355X.

Workload: the variable-density pressure-Poisson system on a 256^3 unit cube,
16 seeded bubbles (radius 0.03-0.12, density 1 inside, 1000 outside), Neumann
walls, right-hand side = divergence of a seeded uniform(-1,1) face-velocity
field shifted to zero mean, x0 = 0.  A step is one complete solve to
||b - A x||_2 <= 1e-8 ||b||_2 by CG preconditioned with the additive-correction
multigrid cycle (2+2 damped Jacobi, coarsest 8^3 exact FP64; on levels 2-4 the
coarse correction is two flexible-CG steps, the K-cycle, on the levels of
16^3 to 64^3 cells; --kcycle 0 gives the V-cycle): FP32 CG with FP64 residual updates (the value), and the plain FP64
solve beside it; the V-cycle's two figures are reported too.  Matrix and
right-hand side are resident in HBM before the timed region.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256] [--no-cpu] [--no-ref-cycle]
                    [--kcycle -1]

`--case batch` solves 96 seeded systems of uneven sizes as one batch and
beside it one after another.  `--case 512` and `--case cyl` print the figures
of the two other single-system
workloads instead (512^3 with 64 bubbles to 1e-10; 384x384x768 z-stretched
grid with a solid cylinder and 128 bubbles to 1e-9): solve times for FP64,
restarted refinement (FP32 inner to 1e-4) and reliable updates, the final
relative residuals, and the error against an FP64 solve to 1e-12.
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


def pressure_step(n, steps, kcycle=0, cfac=2.0):
    """One time step's pressure phase on the device, as a two-phase code runs
    it every step: Pressure::discretize from the density, update_rhs from the
    face velocities, the mixed solve (hierarchy rebuilt for the new matrix)
    and Momentum::project; the divergence left after the projection."""
    rng = np.random.default_rng(7)
    xc = (np.arange(n) + 0.5) / n
    X, Y, Z = np.meshgrid(xc, xc, xc, indexing="ij", sparse=True)
    rho = np.full((n, n, n), 1000.0)
    for _ in range(16):
        c, r = rng.uniform(0.15, 0.85, 3), rng.uniform(0.03, 0.12)
        rho[(X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2 < r * r] = 1.0
    rho = np.pad(rho, 1, mode="edge").ravel()
    u = rng.uniform(-1.0, 1.0, (n + 1, n, n))
    v = rng.uniform(-1.0, 1.0, (n, n + 1, n))
    w = rng.uniform(-1.0, 1.0, (n, n, n + 1))
    u[[0, n]] = 0.0
    v[:, [0, n]] = 0.0
    w[:, :, [0, n]] = 0.0
    h = np.full(n, 1.0 / n)
    met = (h, h, h)
    dr = psb.DeviceArray.from_host(rho)
    du0, dv0, dw0 = (psb.DeviceArray.from_host(a.ravel()) for a in (u, v, w))
    du, dv, dw = psb.DeviceArray(u.size), psb.DeviceArray(v.size), psb.DeviceArray(w.size)
    db, dx, ddiv = psb.DeviceArray(n ** 3), psb.DeviceArray(n ** 3), psb.DeviceArray(n ** 3)
    s = psb.Solver(n, n, n)
    stop = dict(precond=psb.PSB_PREC_AC, norm=psb.PSB_NORM_L2REL, res_rat=1e-8, res_tol=0.0,
                max_iter=1000, precision=psb.PSB_MIXED, inner_rtol=1e-2, inner_adapt=2,
                ac_store=1, ac_kcycle=kcycle, ac_cfac=cfac)
    t = {"assemble": 0.0, "rhs": 0.0, "solve": 0.0, "project": 0.0}
    res = None
    for step in range(steps + 1):
        for src, dst in ((du0, du), (dv0, dv), (dw0, dw)):
            psb.device_copy(dst, src)
        dx.zero()
        psb.device_synchronize()
        t0 = time.perf_counter()
        s.assemble_pressure_device(dr.addr, met, met, met)
        t1 = time.perf_counter()
        s.update_rhs_device(du.addr, dv.addr, dw.addr, db.addr, zero_mean=False)
        t2 = time.perf_counter()
        res = s.solve_device(dx.addr, db.addr, **stop)
        t3 = time.perf_counter()
        s.project_device(dr.addr, dx.addr, du.addr, dv.addr, dw.addr)
        t4 = time.perf_counter()
        if step > 0:  # the first step is a warm-up
            for k, a, b in (("assemble", t0, t1), ("rhs", t1, t2), ("solve", t2, t3),
                            ("project", t3, t4)):
                t[k] += (b - a) / steps
    s.update_rhs_device(du.addr, dv.addr, dw.addr, ddiv.addr, zero_mean=False)
    div, b = ddiv.download(), db.download()
    proj_bytes = 64.0 * n ** 3  # p, rho, u/v/w read and written (FP64)
    ms_p = t["project"] * 1e3
    return {"grid": [n, n, n], "steps": steps,
            "assemble_ms": round(t["assemble"] * 1e3, 3), "rhs_ms": round(t["rhs"] * 1e3, 3),
            "solve_ms": round(t["solve"] * 1e3, 3), "project_ms": round(t["project"] * 1e3, 3),
            "total_ms": round(sum(t.values()) * 1e3, 3),
            "project_GBs_incl_call": round(proj_bytes / (ms_p * 1e-3) / 1e9, 1),
            "iterations": res["iterations"],
            "divergence_after_projection_rel": float(np.linalg.norm(div) / np.linalg.norm(b))}


def measured_traffic(n, kind):
    """HBM bytes per launch of the fine-level kernel `kind` at 256^3 from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summaries (calibrated against
    a plain copy): a line "traffic_bytes_per_launch kind=<kind> <bytes>" in a
    profiles/*hbm_traffic*.md file; None where none was measured."""
    pdir = os.path.join(ROOT, "profiles")
    if n != 256 or not os.path.isdir(pdir):
        return None
    for f in sorted(os.listdir(pdir)):
        if "hbm_traffic" not in f:
            continue
        m = re.search(r"traffic_bytes_per_launch kind=%s (\d+)" % re.escape(kind),
                      open(os.path.join(pdir, f)).read())
        if m:
            return int(m.group(1))
    return None


def ref_driver(mode, n, limit):
    """Run the reference build's driver (oracle/_ref) on the bench case; the
    solve time it prints, or None."""
    drv = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
    if not os.path.exists(drv):
        return None
    args = [drv, mode, str(n), str(CASE["nsph"]), repr(CASE["rmin"]), repr(CASE["rmax"]),
            repr(CASE["rho_in"]), repr(CASE["rho_out"]), str(CASE["seed_sph"]),
            str(CASE["seed_vel"]), str(limit)]
    out = subprocess.run(args, capture_output=True, text=True, timeout=900).stdout
    m = re.search(r"solve_s=([0-9.eE+-]+)", out)
    return float(m.group(1)) if m else None


def host_cpu():
    """CPU model and the cores this process may use (the host the reference's
    CPU sample runs on)."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = os.cpu_count()
    return {"cpu_model": model, "nproc": avail, "cpu_count": os.cpu_count()}


def in_solve_kernels(s, dx, db, stop):
    """One solve with the same settings, launched kernel by kernel, the
    finest level's kernels bracketed by HIP events on the solver's stream:
    per kind the launches, average duration, algorithmic bytes per launch,
    GB/s and fraction of the HBM peak; plus the eager solve's total."""
    dx.zero()
    res, st, solve_ms = s.profile_solve_device(dx.addr, db.addr, **stop)
    out = {}
    for k, v in st.items():
        if not v["launches"]:
            continue
        avg = v["ms"] / v["launches"]
        gbs = v["bytes"] / (avg * 1e-3) / 1e9
        out[k] = {"launches": v["launches"], "avg_us": round(avg * 1e3, 2),
                  "total_ms": round(v["ms"], 3), "bytes_per_launch": v["bytes"],
                  "GBs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
    fine = sum(v["ms"] for v in st.values())
    return {"kernels": out, "eager_solve_ms": round(solve_ms, 3),
            "fine_level_ms": round(fine, 3),
            "coarse_levels_reductions_gaps_ms": round(solve_ms - fine, 3),
            "iterations": res["iterations"]}


def cpu_baseline(n, cg_iters_full, cg_sample):
    """The reference's own solvers, built from its sources into oracle/_ref,
    on one host core, on bounded samples of the same 256^3 system:
      - CG with the diagonal preconditioner (the baseline `value`, the
        reference method that reaches 1e-8 here): `cg_sample` iterations,
        scaled by the iterations it needs to 1e-8 (counted by the GPU's
        parity-tested CG-di in the same run);
      - one cycle of its production multigrid, AC::vcycle with CG smoothers
        (reported beside it: on this system it stops after a few cycles,
        its residual growing again, far above 1e-8)."""
    t_cg = ref_driver("bench-cg", n, cg_sample)
    if t_cg is None:
        return None
    s_iter = t_cg / cg_sample
    out = {"value": 1.0 / (s_iter * cg_iters_full), "unit": "solves/s", "cores": 1,
           "kind": "reference", **host_cpu(),
           "sample": f"reference CG-di (Src/Solver/Linear/Krylov/cg.cpp) {cg_sample} iterations "
                     f"at {n}^3, {s_iter:.3f} s/iteration on 1 core, scaled by the "
                     f"{cg_iters_full} iterations CG-di needs on this system",
           "s_per_iteration": s_iter, "iterations_full": cg_iters_full}
    t_cyc = ref_driver("bench-ac", n, 1)
    if t_cyc is not None:
        out["ac_vcycle_s_per_cycle"] = t_cyc
    return out


EXTRA = {
    "512": dict(grid=(512, 512, 512), gen=dict(nsph=64, rmin=0.02, rmax=0.08, rho_in=1.0,
                                                rho_out=1000.0, seed_sph=3, seed_vel=1003),
                tol=1e-10,
                desc="pressure Poisson 512^3, 64 bubbles r 0.02-0.08, density ratio 1000, "
                     "Neumann walls; AC-PCG to 1e-10"),
    "cyl": dict(grid=(384, 384, 768), gen=dict(nsph=128, rmin=0.01, rmax=0.04, rho_in=1.0,
                                                rho_out=1000.0, seed_sph=5, seed_vel=1005,
                                                stretch_z=2.5, lz=2.0, cylinder=True, cyl_d=0.2),
                tol=1e-9, stop=dict(ac_coarsen=1, ac_cfac=1.0),
                desc="pressure Poisson 384x384x768 on [0,1]x[0,1]x[0,2], tanh z-stretching "
                     "2.5, solid cylinder d 0.2 along y, 128 bubbles r 0.01-0.04, density "
                     "ratio 1000; AC-PCG to 1e-9 (levels halve x and y before z, coarse "
                     "correction factor 1)"),
}


def drop_case(a):
    """The plain reference method at its own size: 64^3, one centred drop,
    CG-di in FP64 to 1e-8 on the GPU, and the reference's CG::solve (its own
    build) on one host core on the same system; beside it CG with the
    red-black Gauss-Seidel preconditioner on the GPU and the reference's CG
    with Prec::ic2 (which it replaces) on one core."""
    n = 64
    d = psb.generate(n, n, n, 1, 0.25, 0.25, 1000.0, 1.0, 0, 1001, centre=(0.5, 0.5, 0.5))
    s = psb.Solver(n, n, n)
    s.set_matrix(d)
    s.set_geometry(d["dxc"], d["dyc"], d["dzc"])
    db = psb.DeviceArray.from_host(d["rhs"])
    dx = psb.DeviceArray(n ** 3)
    t, r = timed_solves(s, dx, db, a.steps, a.warmup, max_iter=1000, res_rat=1e-8,
                        graph_chunk=32)
    out = {"workload": "64^3 unit cube, centred drop r 0.25 (density 1000 in 1), Neumann walls; "
                       "CG with the diagonal preconditioner (the reference method), FP64, to "
                       "1e-8 (reference norm)",
           "gpu_ms_per_solve": round(t * 1e3, 3), "iterations": r["iterations"],
           "rel_residual": r["rel_l2"]}
    t2, r2 = timed_solves(s, dx, db, a.steps, a.warmup, max_iter=1000, res_rat=1e-8,
                          graph_chunk=32, precond=psb.PSB_PREC_SGS)
    out["sgs_gpu_ms_per_solve"] = round(t2 * 1e3, 3)
    out["sgs_iterations"] = r2["iterations"]
    out["sgs_rel_residual"] = r2["rel_l2"]
    drv = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
    if os.path.exists(drv) and not a.no_cpu:
        for mode, key, tg in (("bench-cg", "reference_cpu_s_per_solve", t),
                              ("bench-ic2", "reference_ic2_cpu_s_per_solve", t2)):
            args = [drv, mode, str(n), "1", "0.25", "0.25", "1000.0", "1.0", "0", "1001", "1000",
                    "0.5", "0.5", "0.5"]
            o = subprocess.run(args, capture_output=True, text=True, timeout=600).stdout
            m = re.search(r"solve_s=([0-9.eE+-]+)", o)
            it = re.search(r"iterations = (\d+)", o)
            if m:
                out[key] = float(m.group(1))
                out[key.replace("_s_per_solve", "_over_gpu")] = round(float(m.group(1)) / tg, 1)
            if it and mode == "bench-ic2":
                out["reference_ic2_iterations"] = int(it.group(1))
    print(json.dumps(out), flush=True)


def extra_case(a):
    """Figures for the 512^3 and the stretched-cylinder workloads."""
    E = EXTRA[a.case]
    nx, ny, nz = E["grid"]
    log(f"[bench] generating {a.case} case {nx}x{ny}x{nz}")
    d = psb.generate(nx, ny, nz, **E["gen"])
    s = psb.Solver(nx, ny, nz)
    if d["off"].any():
        s.set_active(d["off"])
    s.set_matrix(d)
    s.set_geometry(d["dxc"], d["dyc"], d["dzc"])
    db = psb.DeviceArray.from_host(d["rhs"])
    dx = psb.DeviceArray(nx * ny * nz)
    del d
    common = dict(precond=psb.PSB_PREC_AC, norm=psb.PSB_NORM_L2REL, res_rat=E["tol"],
                  res_tol=0.0, max_iter=2000, **E.get("stop", {}))
    runs = {"fp64": dict(precision=psb.PSB_FP64),
            "mixed_restarted_1e-4": dict(precision=psb.PSB_MIXED, inner_rtol=1e-4,
                                         inner_adapt=0, max_outer=60),
            "mixed_reliable_1e-2": dict(precision=psb.PSB_MIXED, inner_rtol=1e-2,
                                        inner_adapt=2),
            "mixed_reliable_1e-2_fp16_smoother": dict(precision=psb.PSB_MIXED, inner_rtol=1e-2,
                                                      inner_adapt=2, ac_store=1),
            "mixed_reliable_1e-2_fp16_smoother_and_cg": dict(precision=psb.PSB_MIXED,
                                                             inner_rtol=1e-2, inner_adapt=2,
                                                             ac_store=2)}
    # the K-cycle as the headline runs it (-1: levels chosen from the hierarchy)
    km = a.kcycle
    runs["fp64_kcycle"] = dict(precision=psb.PSB_FP64, ac_kcycle=km)
    runs["mixed_reliable_1e-2_fp16_smoother_kcycle"] = dict(
        precision=psb.PSB_MIXED, inner_rtol=1e-2, inner_adapt=2, ac_store=1, ac_kcycle=km)
    dx.zero()
    r12 = s.solve_device(dx.addr, db.addr, precision=psb.PSB_FP64,
                         **dict(common, res_rat=1e-12, max_iter=4000))
    x_ref = dx.download()
    x_ref -= x_ref.mean()
    log(f"[bench] fp64 to 1e-12 {r12}")
    out = {"workload": E["desc"], "grid": list(E["grid"]), "tolerance": E["tol"], "kcycle": km,
           "fp64_1e-12": {"iterations": r12["iterations"], "rel_l2": r12["rel_l2"]}}
    for name, kw in runs.items():
        t, r = timed_solves(s, dx, db, a.steps, a.warmup, **dict(common, **kw))
        x = dx.download()
        x -= x.mean()
        err = float(np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref))
        log(f"[bench] {name} {t * 1e3:.2f} ms/solve {r} err {err:.2e}")
        out[name] = {"ms_per_solve": round(t * 1e3, 3), "solves_per_s": round(1.0 / t, 3),
                     "iterations": r["iterations"], "outer": r["outer"],
                     "status": r["status_name"], "rel_l2": r["rel_l2"],
                     "error_vs_fp64_1e-12": err}
    out["mixed_over_fp64"] = round(out["fp64"]["ms_per_solve"] /
                                   min(out["mixed_reliable_1e-2"]["ms_per_solve"],
                                       out["mixed_reliable_1e-2_fp16_smoother"]["ms_per_solve"]),
                                   3)
    print(json.dumps(out), flush=True)


def batch_case(a):
    """96 systems of seeded sizes solved as one batch, against the same
    systems solved one after another with the same method (same contexts)."""
    cases = psb.batch_cases(4, a.nsys)
    shapes = [(c["nx"], c["ny"], c["nz"]) for c in cases]
    ncell = sum(x * y * z for x, y, z in shapes)
    log(f"[bench] generating {len(cases)} systems, {ncell / 1e6:.1f} M cells")
    B = psb.Batch(shapes)
    dbs, dxs = [], []
    for q, c in enumerate(cases):
        d = psb.generate(**c)
        B.systems[q].set_matrix(d)
        B.systems[q].set_geometry(d["dxc"], d["dyc"], d["dzc"])
        dbs.append(psb.DeviceArray.from_host(d["rhs"]))
        dxs.append(psb.DeviceArray(d["rhs"].size))
    # the V-cycle (measured faster for this batch: its systems are small
    # enough that the K-cycle's extra small-level launches cost more than its
    # iterations save), and the K-cycle on levels 1 and 2 (by level index)
    # beside it
    kb = 0 if a.kcycle == -1 else a.kcycle
    stop = dict(res_rat=1e-8, max_iter=1000, ac_kcycle=kb)
    px, pb = [x.addr for x in dxs], [b.addr for b in dbs]

    def run_batch(**kw):
        for x in dxs:
            x.zero()
        return B.solve_device(px, pb, **dict(stop, **kw))

    def timed_batch(**kw):
        for _ in range(max(1, a.warmup)):
            res, rc = run_batch(**kw)
        psb.device_synchronize()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            res, rc = run_batch(**kw)
        psb.device_synchronize()
        return (time.perf_counter() - t0) / a.steps, res, rc

    log("[bench] batched solves (plan build and autotuning in the warm-up)")
    tk, resk, _ = timed_batch(ac_kcycle=-1)
    itk = [r["iterations"] for r in resk]
    log(f"[bench] batch K-cycle {tk * 1e3:.2f} ms")
    tb, res, rc = timed_batch()
    xb = [x.download() for x in dxs]
    log(f"[bench] batch {tb * 1e3:.2f} ms rc {rc}")
    # the same systems one after another (same contexts, same method; the
    # batch's K-cycle levels by index)
    kw = dict(psb.BATCH_STOP, **stop)
    for q in range(len(cases)):
        dxs[q].zero()
        B.systems[q].solve_device(dxs[q].addr, dbs[q].addr, **kw)
    psb.device_synchronize()
    t0 = time.perf_counter()
    seq = []
    for q in range(len(cases)):
        dxs[q].zero()
        seq.append(B.systems[q].solve_device(dxs[q].addr, dbs[q].addr, **kw))
    psb.device_synchronize()
    ts = time.perf_counter() - t0
    log(f"[bench] sequential {ts * 1e3:.2f} ms")
    # error against FP64 solves to 1e-12
    errs = []
    for q in range(len(cases)):
        dxs[q].zero()
        B.systems[q].solve_device(dxs[q].addr, dbs[q].addr, precond=psb.PSB_PREC_AC,
                                  norm=psb.PSB_NORM_L2REL, res_rat=1e-12, res_tol=0.0,
                                  max_iter=2000)
        xr = dxs[q].download()
        xr -= xr.mean()
        xq = xb[q] - xb[q].mean()
        errs.append(float(np.linalg.norm(xq - xr) / np.linalg.norm(xr)))
    its = [r["iterations"] for r in res]
    out = {"workload": f"{len(cases)} pressure systems, dims from {{32,48,64,96,128}} per "
                       "direction (seed 4), 1-8 bubbles r 0.05-0.2 of the shortest side, "
                       "density ratio 10-1000; FP32 AC-PCG with FP64 reliable updates to 1e-8"
                       + ("" if kb == 0 else f", K-cycle on levels {kb:#b} (bit = level)"),
           "systems": len(cases), "cells": ncell,
           "batched_ms": round(tb * 1e3, 3), "batched_systems_per_s": round(len(cases) / tb, 2),
           "sequential_ms": round(ts * 1e3, 3),
           "batched_over_sequential": round(ts / tb, 3),
           "converged": sum(r["converged"] for r in res), "status": rc,
           "max_rel_l2": max(r["rel_l2"] for r in res),
           "iterations_min_mean_max": [min(its), sum(its) / len(its), max(its)],
           "max_error_vs_fp64_1e-12": max(errs),
           "plan_level_class_choice": B.plan()}
    out["kcycle_levels_1_2"] = {"batched_ms": round(tk * 1e3, 3),
                                "iterations_min_mean_max": [min(itk), sum(itk) / len(itk),
                                                            max(itk)],
                                "converged": sum(r["converged"] for r in resk)}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ref-cycle", action="store_true",
                    help="leave out the reference AC::vcycle timing")
    ap.add_argument("--cpu-iters", type=int, default=20)
    ap.add_argument("--graph-chunk", type=int, default=8,
                    help="CG iterations per captured graph of the headline solves (smaller "
                         "values give tracing tools smaller graphs)")
    ap.add_argument("--kcycle", type=lambda v: int(v, 0), default=-1,
                    help="K-cycle levels of the solves (bit l: level l; -1: chosen from the "
                         "hierarchy, the levels of 16^3 to 64^3 cells, or the first four "
                         "coarse levels of a semi-coarsened one; 0: V-cycle)")
    ap.add_argument("--cfac", type=float, default=2.0,
                    help="coarse-correction factor of the V-cycle levels (psb_stop.ac_cfac; "
                         "at 256^3 2.2-3.5 give 15 iterations instead of 16, 4 %% faster, "
                         "but the error left at the 1e-8 residual is then 1.2-1.9e-7 "
                         "against 7.9e-8)")
    ap.add_argument("--nsys", type=int, default=96, help="systems of --case batch")
    ap.add_argument("--case", choices=sorted(EXTRA) + ["batch", "drop64"], default=None,
                    help="figures of another single-system workload instead of the bench line")
    ap.add_argument("--headline-only", action="store_true",
                    help="only the timed FP64 and mixed solves of the bench line and the "
                         "in-solve kernel timing (for a kernel-trace run of the bench method)")
    ap.add_argument("--only", choices=("fp64", "mixed"), default=None,
                    help="with --headline-only: only this one of the two solves (one "
                         "precision per traced process)")
    ap.add_argument("--profile-kernels", action="store_true",
                    help="only launch the roofline kernels (FP32 Jacobi sweeps, FP32/FP64 "
                         "SpMV, FP32 CG p/q and x/r steps) and a plain copy, for "
                         "hardware-counter runs")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported: the solver has "
              "no multi-GPU path (it would shard by 3-D blocks with a one-cell halo exchange "
              "and all-reduced scalars)", file=sys.stderr)
        sys.exit(2)
    if a.case == "batch":
        batch_case(a)
        return
    if a.case == "drop64":
        drop_case(a)
        return
    if a.case:
        extra_case(a)
        return
    n = a.n
    log(f"[bench] generating {n}^3 case")
    d = psb.generate(n, n, n, **CASE)
    s = psb.Solver(n, n, n)
    s.set_matrix(d)
    s.set_geometry(d["dxc"], d["dyc"], d["dzc"])
    if a.profile_kernels:
        # the mixed solve's hierarchy (FP16 smoother copies) for the Jacobi sweep
        dx0 = psb.DeviceArray(n ** 3)
        db0 = psb.DeviceArray.from_host(d["rhs"])
        s.solve_device(dx0.addr, db0.addr, precond=psb.PSB_PREC_AC, precision=psb.PSB_MIXED,
                       inner_adapt=2, inner_rtol=1e-2, ac_store=1, ac_kcycle=a.kcycle,
                       max_iter=1)
        for kern, name in ((2, "jacobi_fp32"), (1, "spmv_fp32"), (0, "spmv_fp64"),
                           (6, "cg_pq_fp32"), (7, "sweep_prol_fp32"), (8, "sweep_last_fp32"),
                           (9, "cg_xr_fp32"), (3, "copy_fp32")):
            ms, by = s.time_kernel(kern, reps=10)
            log(f"[bench] {name}: {ms * 1e3:.1f} us/launch, {by / 1e6:.1f} MB algorithmic, "
                f"{by / ms / 1e6:.0f} GB/s")
        return
    db = psb.DeviceArray.from_host(d["rhs"])
    dx = psb.DeviceArray(n ** 3)
    common = dict(precond=psb.PSB_PREC_AC, norm=psb.PSB_NORM_L2REL, res_rat=1e-8, res_tol=0.0,
                  max_iter=1000, ac_kcycle=a.kcycle, ac_cfac=a.cfac, graph_chunk=a.graph_chunk)
    mixed = dict(precision=psb.PSB_MIXED, inner_rtol=1e-2, inner_adapt=2, ac_store=1)
    if a.headline_only and a.only:
        kw = dict(common, **(mixed if a.only == "mixed" else dict(precision=psb.PSB_FP64)))
        t, r = timed_solves(s, dx, db, a.steps, a.warmup, **kw)
        log(f"[bench] {a.only} {t * 1e3:.2f} ms/solve {r}")
        prof = in_solve_kernels(s, dx, db, kw)
        log(f"[bench] in-solve {a.only} {prof}")
        print(json.dumps({a.only + "_ms_per_solve": round(t * 1e3, 3),
                          "iterations": r["iterations"], "in_solve": prof}), flush=True)
        return
    log(f"[bench] FP64 AC-PCG, K-cycle levels {a.kcycle:#b}")
    t64, r64 = timed_solves(s, dx, db, a.steps, a.warmup, precision=psb.PSB_FP64, **common)
    log(f"[bench] fp64 {t64 * 1e3:.2f} ms/solve {r64}")
    x_fp64 = dx.download()
    # the same two solves with the plain V-cycle, for comparison
    vc = dict(common, ac_kcycle=0)
    if not a.headline_only:
        t64v, r64v = timed_solves(s, dx, db, a.steps, a.warmup, precision=psb.PSB_FP64, **vc)
        tmxv, rmxv = timed_solves(s, dx, db, a.steps, a.warmup, **dict(vc, **mixed))
        log(f"[bench] V-cycle: fp64 {t64v * 1e3:.2f} ms ({r64v['iterations']}), mixed "
            f"{tmxv * 1e3:.2f} ms ({rmxv['iterations']})")
    log("[bench] mixed AC-PCG, FP32 iterations with FP64 reliable updates")
    tmx, rmx = timed_solves(s, dx, db, a.steps, a.warmup, **dict(common, **mixed))
    log(f"[bench] mixed {tmx * 1e3:.2f} ms/solve {rmx}")
    x_mixed = dx.download()
    if a.headline_only:
        prof_mx = in_solve_kernels(s, dx, db, dict(common, **mixed))
        prof_64 = in_solve_kernels(s, dx, db, dict(common, precision=psb.PSB_FP64))
        log(f"[bench] in-solve mixed {prof_mx}")
        print(json.dumps({"mixed_ms_per_solve": round(tmx * 1e3, 3),
                          "fp64_ms_per_solve": round(t64 * 1e3, 3),
                          "mixed_over_fp64": round(t64 / tmx, 3),
                          "iterations": {"fp64": r64["iterations"], "mixed": rmx["iterations"]},
                          "in_solve_mixed": prof_mx, "in_solve_fp64": prof_64}), flush=True)
        return
    # the method as first specified: FP32 smoother matrices (no FP16 copy)
    mixed32 = dict(mixed, ac_store=0)
    tm32, rm32 = timed_solves(s, dx, db, a.steps, a.warmup, **dict(common, **mixed32))
    log(f"[bench] mixed, FP32 smoother {tm32 * 1e3:.2f} ms/solve {rm32}")
    x_m32 = dx.download()
    # FP64 reference solution to 1e-12 for the error figure (untimed)
    dx.zero()
    r12 = s.solve_device(dx.addr, db.addr, precision=psb.PSB_FP64, **dict(common, res_rat=1e-12))
    x_ref = dx.download()
    xr = x_ref - x_ref.mean()

    def err_of(x):
        x = x - x.mean()
        return float(np.linalg.norm(x - xr) / np.linalg.norm(xr))

    err, err32 = err_of(x_mixed), err_of(x_m32)
    # the same figure for the FP64 solve to the same tolerance: what a 1e-8
    # residual allows on this system, whatever the precision
    err64 = err_of(x_fp64)
    # the residual tolerance each path needs for a given solution error
    # (untimed): mixed <= 1e-7, FP64 <= 1e-8
    tol_sweep = {"mixed": [], "fp64": []}
    for tol in (1e-8, 5e-9, 2e-9, 1e-9, 5e-10, 2e-10, 1e-10):
        for name, kw in (("mixed", mixed), ("fp64", dict(precision=psb.PSB_FP64))):
            dx.zero()
            rr = s.solve_device(dx.addr, db.addr, **dict(common, res_rat=tol, **kw))
            tol_sweep[name].append({"res_rat": tol, "iterations": rr["iterations"],
                                    "rel_l2": rr["rel_l2"], "error": err_of(dx.download())})
    tol_for = {name: next((t["res_rat"] for t in tol_sweep[name] if t["error"] <= lim), None)
               for name, lim in (("mixed", 1e-7), ("fp64", 1e-8))}
    log(f"[bench] tolerance for the error bounds {tol_for}")
    # per-kernel durations inside the solve (the finest level's kernels
    # bracketed by HIP events in an eager run of the same solve)
    prof_mx = in_solve_kernels(s, dx, db, dict(common, **mixed))
    prof_64 = in_solve_kernels(s, dx, db, dict(common, precision=psb.PSB_FP64))
    log(f"[bench] in-solve mixed {prof_mx}")
    log(f"[bench] in-solve fp64 {prof_64}")
    # the dominant kernel: the fine-level kind with the largest time per solve
    kern = max(prof_mx["kernels"].items(), key=lambda kv: kv[1]["total_ms"])
    names = {"cg_pq": "k_cg_pq_lds<float,4,SYM,-,flux> (p = z + beta p, q = A p, p.q; FP32 "
                      "matrix, operand plane staged in LDS)",
             "sweep_prol": "k_jacobi_lds<float,4,PROL,-,_Float16,flux> (first post-sweep with "
                           "the prolongation fused; FP16 matrix copy, LDS-staged)",
             "sweep": "k_jacobi_march<float,4,-,-,_Float16,flux> (damped Jacobi sweep, FP16 "
                      "matrix copy, marching)",
             "sweep_last": "k_jacobi_march<float,4,-,RHO,_Float16,flux> (last post-sweep with "
                           "the rho partials, marching)",
             "restrict": "k_restrict_march<float,4,...> (residual block sums, coarse first "
                         "sweep fused)",
             "cg_xr": "k_cg_xr<float,4> (x, r update, norms, next first sweep)"}
    # isolated back-to-back launches of the fine-level sweep and SpMVs
    ms_j, by_j = s.time_kernel(2, reps=50)
    ms_s32, by_s32 = s.time_kernel(1, reps=50)
    ms_s64, by_s64 = s.time_kernel(0, reps=50)
    gbs = lambda by, ms: by / (ms * 1e-3) / 1e9
    kv = kern[1]
    roof = {"bound": "hbm", "kernel": names.get(kern[0], kern[0]), "kind": kern[0],
            "achieved": kv["GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": kv["frac"], "traffic": measured_traffic(n, kern[0]),
            "avg_launch_ms": kv["avg_us"] * 1e-3, "bytes_per_launch": kv["bytes_per_launch"],
            "measured": "in-solve: HIP events around each launch on the solver's stream in an "
                        "eager run of the headline solve",
            "in_solve_fine_kernels": prof_mx["kernels"],
            "isolated": {"jacobi_fp16_GBs": round(gbs(by_j, ms_j), 1),
                         "spmv_fp32_GBs": round(gbs(by_s32, ms_s32), 1),
                         "spmv_fp64_GBs": round(gbs(by_s64, ms_s64), 1),
                         "note": "50 back-to-back launches on the same arrays"}}
    # the reference's own multigrid (AC::vcycle) on the GPU: cycles and time
    ref_cycle = None
    if not a.no_ref_cycle:
        log("[bench] reference AC::vcycle on the GPU")
        dx.zero()
        rv = s.ac_vcycle_device(dx.addr, db.addr, res_rat=1e-8, max_cycles=50)
        psb.device_synchronize()
        t0 = time.perf_counter()
        dx.zero()
        rv = s.ac_vcycle_device(dx.addr, db.addr, res_rat=1e-8, max_cycles=50)
        psb.device_synchronize()
        t_vc = time.perf_counter() - t0
        log(f"[bench] AC::vcycle {t_vc * 1e3:.1f} ms {rv}")
        ref_cycle = {"ms": round(t_vc * 1e3, 2), "cycles": rv["iterations"],
                     "status": rv["status_name"], "rel_residual": rv["rel_l2"]}
    log("[bench] pressure phase of a time step (assembly, rhs, solve, projection)")
    step = pressure_step(n, 3, a.kcycle, a.cfac)
    log(f"[bench] pressure step {step}")
    cpu = None
    if not a.no_cpu:
        log("[bench] FP64 CG-di iteration count for the CPU baseline scale")
        dx.zero()
        rdi = s.solve_device(dx.addr, db.addr, precision=psb.PSB_FP64, norm=psb.PSB_NORM_L2REL,
                             res_rat=1e-8, res_tol=0.0, max_iter=100000, graph_chunk=64)
        log(f"[bench] CG-di {rdi}")
        log("[bench] reference CPU samples")
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
        "dtype": "fp32 iterations / fp64 residual updates (fp16 smoother matrix)",
        "data": "synthetic (seeded bubbles, seeded face velocities)",
        "config": {"workload": f"pressure Poisson {n}^3, 16 bubbles r 0.03-0.12, density "
                               "ratio 1000, Neumann walls; AC-preconditioned CG (2+2 damped "
                               f"Jacobi, omega 0.9, coarse-correction factor {a.cfac:g}, 2x "
                               "coarsening to 8^3, exact FP64 coarsest, two "
                               "flexible-CG steps as the coarse correction of "
                               + ("the levels of 64^3 to 16^3 cells" if a.kcycle == -1 else
                                  f"levels {a.kcycle:#b} (bit = level)")
                               + ") to 1e-8; FP32 CG with an FP64 true-residual update per "
                               "1e-2 drop, FP16 copy of the level matrices in the smoother",
                   "grid": [n, n, n], "parallelism": "single GPU"},
        "fp64_solves_per_s": round(1.0 / t64, 3),
        "fp64_ms_per_solve": round(t64 * 1e3, 3),
        "mixed_over_fp64": round(t64 / tmx, 3),
        "iterations": {"fp64": r64["iterations"], "mixed": rmx["iterations"],
                       "mixed_fp64_updates": rmx["outer"]},
        "vcycle": {"fp64_ms_per_solve": round(t64v * 1e3, 3),
                   "mixed_ms_per_solve": round(tmxv * 1e3, 3),
                   "iterations": {"fp64": r64v["iterations"], "mixed": rmxv["iterations"]},
                   "mixed_over_fp64": round(t64v / tmxv, 3)},
        "fp32_smoother": {"ms_per_solve": round(tm32 * 1e3, 3),
                          "solves_per_s": round(1.0 / tm32, 3),
                          "mixed_over_fp64": round(t64 / tm32, 3),
                          "iterations": rm32["iterations"], "rel_l2": rm32["rel_l2"],
                          "error_vs_fp64_1e-12": err32},
        "rel_residual": {"fp64": r64["rel_l2"], "mixed": rmx["rel_l2"]},
        "error_vs_fp64_1e-12": err,
        "fp64_solve_error_vs_fp64_1e-12": err64,
        "tolerance_for_error": {"mixed_error_le_1e-7": tol_for["mixed"],
                                "fp64_error_le_1e-8": tol_for["fp64"], "sweep": tol_sweep},
        "in_solve": {"mixed": prof_mx, "fp64": prof_64},
        "reference_ac_vcycle_gpu": ref_cycle,
        "pressure_step_gpu": step,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
