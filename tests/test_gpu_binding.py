"""The reference-side binding, GpuKrylov (include/psb_krylov.hpp), inside the
reference's own callers of Linear::solve, against the same callers with the
reference's CG on the CPU, on one seeded 16^3 bubble system (density ratio
1000): oracle/_ref/gk_driver links the reference's objects (oracle/ref.mk)
with libpsb.so.

- Centered::solve (centered_solve.cpp:9-40) of the Pressure equation to
  ResRat 1e-8: the same iteration count (+-1) and solution.
- AC::vcycle (additive.h:39-41), whose levels all hold the one solver object
  (centered.cpp:50) and call it on their own coarse matrices
  (additive_cycle_functions.cpp:89-107): the binding keeps one library
  context per level shape, and the cycle count and solution match."""
import os
import re
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRV = os.path.join(ROOT, "oracle", "_ref", "gk_driver")
CASE = ["16", "2", "0.1", "0.2", "1.0", "1000.0", "3", "1003"]
NUM = r"([-+0-9.eE]+|nan|inf)"


def run(mode, tmp):
    if not os.path.exists(DRV):
        pytest.skip("oracle/_ref/gk_driver is not built (the reference sources were absent)")
    print(f"[binding] gk_driver {mode} {' '.join(CASE)}", flush=True)
    r = subprocess.run([DRV, mode, str(tmp)] + CASE, capture_output=True, text=True,
                       timeout=240)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    phi = {t: np.fromfile(os.path.join(tmp, t + "_phi.f64")) for t in ("cpu", "gpu")}
    return r.stdout, phi


def demeaned_rel(a, b):
    a, b = a - a.mean(), b - b.mean()
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def contexts(out):
    m = re.search(r"CONTEXTS (\d+) STATUS (-?\d+)", out)
    return int(m.group(1)), int(m.group(2))


def test_centered_solve_with_the_binding(tmp_path):
    out, phi = run("centered", tmp_path)
    logs = [(float(m.group(1)), float(m.group(2)), int(m.group(3)))
            for m in re.finditer(r"centered, residual = " + NUM + ", ratio = " + NUM +
                                 r", iterations = (\d+)", out)]
    assert len(logs) == 2, out[-2000:]
    (rc, qc, ic), (rg, qg, ig) = logs
    assert abs(ic - ig) <= 1, (ic, ig)
    assert qc < 1e-8 and qg < 1e-8
    assert demeaned_rel(phi["gpu"], phi["cpu"]) < 1e-6
    assert "GpuKrylov:" not in out
    assert contexts(out) == (1, 0)


def test_ac_vcycle_with_the_binding(tmp_path):
    out, phi = run("ac", tmp_path)
    ret = {m.group(1): (int(m.group(2)), int(m.group(3)))
           for m in re.finditer(r"ACRET (\w+) (\d) (\d+)", out)}
    assert ret["cpu"][0] == 1 and ret["gpu"][0] == 1, ret
    assert abs(ret["cpu"][1] - ret["gpu"][1]) <= 1, ret
    assert demeaned_rel(phi["gpu"], phi["cpu"]) < 1e-5
    assert "GpuKrylov:" not in out
    # 16^3 -> 8^3 -> 4^3 (the reference's coarsening): one context per level;
    # the level solves end at their MaxIter (status 1), as the reference's do
    n, status = contexts(out)
    assert n == 3 and status >= 0, (n, status)
