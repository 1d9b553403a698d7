"""Write the golden fixtures of the pressure-solve path from the reference itself.

Runs oracle/_ref/ref_driver (the reference's own Pressure, CG, BiCGS, Jacobi
and AC classes compiled from /root/reference/Src by oracle/ref.mk) on small
synthetic cases and stores what it computed as tests/golden/<case>.npz:
inputs (grid nodes, metrics, density, matrix, right-hand side), operator
outputs, solver iterates after k iterations, AC level matrices, restriction /
interpolation results and V-cycle iterates, plus the solver log lines (17
significant digits) parsed into arrays.

    python tests/golden/make_golden.py [cases] [cycles] [cycles-bicgs] [cycles-nonsym] [flex] [prec] [ghost] [bnd] [case names]   # needs oracle/_ref built

"cycles" writes <case>_cycles.npz: the reference's AC::cycle with the F, W
and full-cycle shapes and its controls (ref_driver cycles), phi after each
run and the run's residual log.  "cycles-bicgs" writes <case>_cycles_bicgs.npz:
six of those runs with BiCGS (Prec::di) as every level's solver, the solver
the reference's AC hands its levels for a non-symmetric equation.
"bnd" writes <case>_bnd.npz: the matrix, boundary source fbnd and right-hand
side with Dirichlet and non-zero Neumann pressure boundaries.
"ghost" writes <case>_ghost.npz: the reference's Pressure::ghost source
(fext) for the case's colour function and a seeded curvature field.
"cycles-nonsym" writes <case>_cycles_nonsym.npz: the same runs on the case's
matrix made non-symmetric by the driver (east and top couplings scaled
per cell, a time term on the diagonal), with that matrix (N_A_*).  "flex" writes <case>_flex.npz: the same for
AC::flexcycle on the level settings where its recursion ends.  "prec" writes <case>_prec.npz: CG with the
reference's incomplete-Cholesky preconditioners IC0 and IC2 (ref_driver
prec), x after 1 and 5 iterations and at ResRat 1e-8, and the final log
lines.

Only the outputs travel; the reference does not.
"""
import json
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_driver")

# name: (nx, ny, nz, nsph, rmin, rmax, rho_in, rho_out, seed_sph, seed_vel, centre)
CASES = {
    "drop12": (12, 12, 12, 1, 0.25, 0.25, 1000.0, 1.0, 1, 1001, (0.5, 0.5, 0.5)),
    "bubbles16x12x8": (16, 12, 8, 3, 0.1, 0.25, 1.0, 1000.0, 5, 1005, None),
    "bubbles16": (16, 16, 16, 4, 0.08, 0.2, 1.0, 1000.0, 2, 1002, None),
    "stretch12x10x8": (12, 10, 8, 3, 0.1, 0.25, 1.0, 1000.0, 9, 1009, None),
}

# cases on a non-uniform grid (ref_driver case-s): the reference's Grid1D
# stretched from a first to a last cell size per direction (d1x, dnx, d1y,
# dny, d1z, dnz in units of h = 1 / max(nx, ny, nz)), so that the distance
# between cell centres differs from the cell size (dxw != dxc)
STRETCH = {"stretch12x10x8": (0.5, 2.0, 2.0, 0.7, 1.0, 3.0)}

# cases of the incomplete-Cholesky runs (ref_driver prec)
PREC_CASES = ["drop12", "bubbles16x12x8", "bubbles16"]

KEEP_F64 = ["xn", "yn", "zn", "dxc", "dxw", "dxe", "dyc", "dys", "dyn", "dzc", "dzb", "dzt",
            "spheres", "rho", "A_c", "A_w", "A_e", "A_s", "A_n", "A_b", "A_t", "A_ci", "dV",
            "b", "x_rand", "Ax_rand", "r_rand", "cg_x1", "cg_x5", "cg_xfull", "bicgs_x1",
            "bicgs_xfull", "jacobi_x5", "restrict_phi", "restrict_fH", "interp_phiH",
            "interp_phih", "ac_phi1", "ac_phifull"]

NUM = r"([-+0-9.eE]+|nan|inf)"


def parse_log(text):
    info = {"cg": [], "bicgs": [], "jacobi": [], "ret": [], "levels": [], "cycles": [],
            "acret": []}
    for line in text.splitlines():
        m = re.match(r"(cg|bicgs|jacobi), residual = " + NUM + ", ratio = " + NUM +
                     ", iterations = (\d+)", line)
        if m:
            info[m.group(1)].append([float(m.group(2)), float(m.group(3)), int(m.group(4))])
            continue
        m = re.match(r"RET (\w+) (\w+) (\d)", line)
        if m:
            info["ret"].append([m.group(1), m.group(2), int(m.group(3))])
            continue
        m = re.match(r"LEVEL (\d+) (\d+) (\d+) (\d+)", line)
        if m:
            info["levels"].append([int(m.group(i)) for i in range(2, 5)])
            continue
        m = re.match(r"Cycle (\d+); res = " + NUM, line)
        if m:
            info["cycles"].append([int(m.group(1)), float(m.group(2))])
            continue
        m = re.match(r"ACRET (\w+) (\d) (\d+)", line)
        if m:
            info["acret"].append([m.group(1), int(m.group(2)), int(m.group(3))])
        m = re.match(r"Initial res = " + NUM, line)
        if m:
            info.setdefault("initial", []).append(float(m.group(1)))
    return info


def run_case(name, spec):
    nx, ny, nz, nsph, rmin, rmax, rin, rout, ss, sv, centre = spec
    stretch = STRETCH.get(name)
    with tempfile.TemporaryDirectory() as d:
        args = [DRIVER, "case-s" if stretch else "case", d, str(nx), str(ny), str(nz), str(nsph),
                repr(rmin), repr(rmax), repr(rin), repr(rout), str(ss), str(sv)]
        if stretch is not None:
            args += [repr(v) for v in stretch]
        if centre is not None:
            args += [repr(c) for c in centre]
        out = subprocess.run(args, check=True, capture_output=True, text=True).stdout
        arrays = {}
        for k in KEEP_F64:
            arrays[k] = np.fromfile(os.path.join(d, k + ".f64"))
        info = parse_log(out)
        for l in range(1, len(info["levels"])):
            for c in "cwesnbt":
                arrays[f"L{l}_{c}"] = np.fromfile(os.path.join(d, f"L{l}_{c}.f64"))
        for l in range(len(info["levels"])):
            arrays[f"L{l}_aflag"] = np.fromfile(os.path.join(d, f"L{l}_aflag.i32"), dtype=np.int32)
    meta = {"name": name, "shape": [nx, ny, nz], "nsph": nsph, "rmin": rmin, "rmax": rmax,
            "rho_in": rin, "rho_out": rout, "seed_sph": ss, "seed_vel": sv,
            "centre": list(centre) if centre is not None else None,
            "stretch": list(stretch) if stretch is not None else None,
            "cg_snapshots": [1, 5], "bicgs_snapshots": [1], "info": info}
    arrays["meta"] = np.array(json.dumps(meta))
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **arrays)
    return path


# cases of the AC::cycle runs (ref_driver cycles): F, W, full-cycle starts,
# linf control, divergence / min-cycle controls, level tolerances
CYCLE_CASES = ["bubbles16x12x8", "bubbles16"]


def parse_cycles(text):
    """ACRUN header (the run's settings), the cycle log lines, ACRET."""
    runs, cur = [], None
    for line in text.splitlines():
        f = line.split()
        if line.startswith("ACRUN "):
            cur = {"tag": f[1], "kind": int(f[2]), "factor": float(f[3]), "max_cyc": int(f[4]),
                   "use_linf": bool(int(f[5])), "stop_if_div": bool(int(f[6])),
                   "priority_min_cyc": bool(int(f[7])), "min_cyc": int(f[8]),
                   "resrat": float(f[9]), "mv": [int(f[10]), int(f[11]), int(f[12])],
                   "ms": [int(f[13]), int(f[14]), int(f[15])],
                   "initial": None, "cycles": []}
            runs.append(cur)
            continue
        m = re.match(r"Initial res = " + NUM + " ; " + NUM, line)
        if m and cur is not None:
            cur["initial"] = [float(m.group(1)), float(m.group(2))]
            continue
        m = re.match(r"Cycle (\d+); res = " + NUM + " ; " + NUM, line)
        if m and cur is not None:
            cur["cycles"].append([int(m.group(1)), float(m.group(2)), float(m.group(3))])
            continue
        m = re.match(r"ACRET (\w+) (\d) (\d+)", line)
        if m and cur is not None:
            cur["converged"], cur["ncyc"] = bool(int(m.group(2))), int(m.group(3))
    return runs


def run_cycles(name, mode="cycles"):
    """mode "cycles": CG level solves; "cycles-bicgs": BiCGS (Prec::di) as
    every level's solver, a subset of the runs (tags "B..."), written to
    <case>_cycles_bicgs.npz; "cycles-nonsym": those runs on a non-symmetric
    matrix, written with it to <case>_cycles_nonsym.npz."""
    nx, ny, nz, nsph, rmin, rmax, rin, rout, ss, sv, centre = CASES[name]
    with tempfile.TemporaryDirectory() as d:
        args = [DRIVER, mode, d, str(nx), str(ny), str(nz), str(nsph), repr(rmin),
                repr(rmax), repr(rin), repr(rout), str(ss), str(sv)]
        if centre is not None and mode == "cycles":
            args += [repr(c) for c in centre]
        out = subprocess.run(args, check=True, capture_output=True, text=True).stdout
        runs = parse_cycles(out)
        arrays = {r["tag"] + "_phi": np.fromfile(os.path.join(d, r["tag"] + "_phi.f64"))
                  for r in runs}
        if mode == "cycles-nonsym":
            for k in "cwesnbt":
                arrays["N_A_" + k] = np.fromfile(os.path.join(d, "N_A_" + k + ".f64"))
    arrays["meta"] = np.array(json.dumps({"name": name, "runs": runs}))
    path = os.path.join(HERE, name + "_" + mode.replace("-", "_") + ".npz")
    np.savez_compressed(path, **arrays)
    return path


# pressure boundary conditions per side (imin, imax, jmin, jmax, kmin, kmax):
# (type 0 neumann / 1 dirichlet, value) -- ref_driver case-bnd
BND_RUNS = {"bubbles16x12x8": ((1, 0.3), (0, 0.0), (0, 0.0), (0, 0.05), (0, 0.0), (1, -0.2))}


def run_bnd(name):
    """<case>_bnd.npz: the case's matrix, fbnd and right-hand side with the
    pressure boundary conditions of BND_RUNS (Centered::create_system_bnd;
    the metrics, density and face velocities are those of <case>.npz)."""
    nx, ny, nz, nsph, rmin, rmax, rin, rout, ss, sv, centre = CASES[name]
    sides = BND_RUNS[name]
    with tempfile.TemporaryDirectory() as d:
        args = [DRIVER, "case-bnd", d, str(nx), str(ny), str(nz), str(nsph), repr(rmin),
                repr(rmax), repr(rin), repr(rout), str(ss), str(sv)]
        for t, v in sides:
            args += [str(t), repr(v)]
        subprocess.run(args, check=True, capture_output=True, text=True)
        arrays = {k: np.fromfile(os.path.join(d, k + ".f64"))
                  for k in ("A_c", "A_w", "A_e", "A_s", "A_n", "A_b", "A_t", "b", "fbnd", "rho")}
    arrays["types"] = np.array([t for t, _ in sides], dtype=np.int32)
    arrays["values"] = np.array([v for _, v in sides])
    path = os.path.join(HERE, name + "_bnd.npz")
    np.savez_compressed(path, **arrays)
    return path


# Pressure::ghost runs (ref_driver ghost): surface tension and the seed of
# the curvature field
GHOST_RUNS = {"bubbles16x12x8": (0.072, 77), "stretch12x10x8": (0.5, 78)}


def run_ghost(name):
    """<case>_ghost.npz: the colour function and a seeded curvature field
    with their ghost layers, sigma, and the fext that the reference's
    Pressure::ghost adds (ref_driver ghost; the case's metrics and density
    are those of <case>.npz)."""
    nx, ny, nz, nsph, rmin, rmax, rin, rout, ss, sv, centre = CASES[name]
    sigma, seed = GHOST_RUNS[name]
    stretch = STRETCH.get(name)
    with tempfile.TemporaryDirectory() as d:
        args = [DRIVER, "ghost-s" if stretch else "ghost", d, str(nx), str(ny), str(nz),
                str(nsph), repr(rmin), repr(rmax), repr(rin), repr(rout), str(ss), str(sv),
                repr(sigma), str(seed)]
        if stretch is not None:
            args += [repr(v) for v in stretch]
        subprocess.run(args, check=True, capture_output=True, text=True)
        arrays = {k: np.fromfile(os.path.join(d, k + ".f64"))
                  for k in ("c_g", "kappa_g", "fext", "rho", "dxc")}
    arrays["sigma"] = np.array(sigma)
    path = os.path.join(HERE, name + "_ghost.npz")
    np.savez_compressed(path, **arrays)
    return path


# AC::flexcycle runs (ref_driver flex): level-solve ResRat and MaxIter for
# which the reference's recursion ends (it does not on many settings: once a
# level's smoother cannot meet its ResRat it alternates between two levels)
FLEX_RUNS = {
    "bubbles16": [(0.5, 20, (20, 20, 20)), (0.2, 20, (20, 20, 20))],
    "bubbles16x12x8": [(0.5, 20, (10, 10, 10)), (0.2, 20, (20, 20, 20)),
                       (0.2, 20, (10, 10, 10)), (0.1, 20, (20, 20, 20)),
                       (0.1, 20, (10, 10, 10))],
}


def run_flex(name):
    nx, ny, nz, nsph, rmin, rmax, rin, rout, ss, sv, centre = CASES[name]
    runs, arrays = [], {}
    for rr, maxc, mv in FLEX_RUNS[name]:
        with tempfile.TemporaryDirectory() as d:
            args = [DRIVER, "flex", d, str(nx), str(ny), str(nz), str(nsph), repr(rmin),
                    repr(rmax), repr(rin), repr(rout), str(ss), str(sv), repr(rr), str(maxc),
                    *[str(m) for m in mv]]
            out = subprocess.run(args, check=True, capture_output=True, text=True,
                                 timeout=120).stdout
            run = parse_cycles(out)[0]
            run["tag"] = f"Flex_{rr}_{mv[0]}"
            arrays[run["tag"] + "_phi"] = np.fromfile(os.path.join(d, "Flex_phi.f64"))
            runs.append(run)
    arrays["meta"] = np.array(json.dumps({"name": name, "runs": runs}))
    path = os.path.join(HERE, name + "_flex.npz")
    np.savez_compressed(path, **arrays)
    return path


def run_prec(name):
    """CG with the reference's Prec::ic0 / Prec::ic2 (ref_driver prec): x
    after 1 and 5 iterations and at 1e-8, the printed final residuals."""
    nx, ny, nz, nsph, rmin, rmax, rin, rout, ss, sv, centre = CASES[name]
    with tempfile.TemporaryDirectory() as d:
        args = [DRIVER, "prec", d, str(nx), str(ny), str(nz), str(nsph), repr(rmin),
                repr(rmax), repr(rin), repr(rout), str(ss), str(sv)]
        if centre is not None:
            args += [repr(c) for c in centre]
        out = subprocess.run(args, check=True, capture_output=True, text=True).stdout
        runs = {}
        for line in out.splitlines():
            m = re.match(r"(ic0|ic2), residual = " + NUM + ", ratio = " + NUM +
                         r", iterations = (\d+)", line)
            if m:
                runs.setdefault(m.group(1), []).append(
                    [float(m.group(2)), float(m.group(3)), int(m.group(4))])
            m = re.match(r"RET (ic0|ic2) (\w+) (\d)", line)
            if m:
                runs.setdefault(m.group(1) + "_ret", []).append([m.group(2), int(m.group(3))])
        arrays = {f"{t}_x{k}": np.fromfile(os.path.join(d, f"{t}_x{k}.f64"))
                  for t in ("ic0", "ic2") for k in ("1", "5", "full")}
    arrays["meta"] = np.array(json.dumps({"name": name, "snapshots": [1, 5, "full"],
                                          "runs": runs}))
    path = os.path.join(HERE, name + "_prec.npz")
    np.savez_compressed(path, **arrays)
    return path


def main():
    if not os.path.exists(DRIVER):
        sys.exit("oracle/_ref/ref_driver is not built (make -f oracle/ref.mk)")
    what = sys.argv[1:] or ["cases", "cycles", "cycles-bicgs", "cycles-nonsym", "flex", "prec",
                            "ghost", "bnd"]
    only = [w for w in what if w in CASES]  # restrict to the cases named
    if "cases" in what:
        for name, spec in CASES.items():
            if not only or name in only:
                print(run_case(name, spec))
    if "cycles" in what:
        for name in CYCLE_CASES:
            if not only or name in only:
                print(run_cycles(name))
    if "cycles-bicgs" in what:
        for name in CYCLE_CASES:
            if not only or name in only:
                print(run_cycles(name, "cycles-bicgs"))
    if "cycles-nonsym" in what:
        for name in CYCLE_CASES:
            if not only or name in only:
                print(run_cycles(name, "cycles-nonsym"))
    if "bnd" in what:
        for name in BND_RUNS:
            if not only or name in only:
                print(run_bnd(name))
    if "ghost" in what:
        for name in GHOST_RUNS:
            if not only or name in only:
                print(run_ghost(name))
    if "flex" in what:
        for name in FLEX_RUNS:
            if not only or name in only:
                print(run_flex(name))
    if "prec" in what:
        for name in PREC_CASES:
            if not only or name in only:
                print(run_prec(name))


if __name__ == "__main__":
    main()
