"""The reference-side binding (include/psb_krylov.hpp, class GpuKrylov)
compiles against the reference's own headers as a Krylov solver that its
callers accept (Centered and AC take a Linear *, centered.h:81,
additive.h:30).  CPU only; needs the reference sources (skipped where they
are absent)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/Src"
CFG = os.path.join(ROOT, "oracle", "_ref", "cfg")

TU = r"""
#include "Solver/Linear/Krylov/krylov.h"
#include "Solver/Additive/additive.h"
#include "Equation/Centered/Pressure/pressure.h"
#include "psb_krylov.hpp"
#include <type_traits>
static_assert(std::is_base_of<Linear, GpuKrylov>::value, "a Linear solver");
static_assert(std::is_base_of<Krylov, GpuKrylov>::value, "a Krylov solver");
// what a case main writes: the binding in place of CG(d, Prec::di()), handed
// to the Pressure equation and through it to every AC level
void case_main(const Domain &d, const Scalar &p, const Scalar &f, const Vector &uvw,
               Times &t, Matter *m) {
  GpuKrylov solver(d);
  solver.settings().precision = PSB_FP64;
  Linear *as_linear = &solver;
  Pressure pr(p, f, uvw, t, as_linear, m);
  AC multigrid(&pr);
  (void)multigrid;
}
"""


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference sources absent")
def test_binding_compiles_against_reference_headers(tmp_path):
    if not os.path.exists(os.path.join(CFG, "Parallel", "mpi_macros.h")):
        subprocess.check_call(["make", "-s", "-f", "oracle/ref.mk",
                               os.path.join("oracle", "_ref", "cfg", ".stamp")], cwd=ROOT)
    src = tmp_path / "binding.cpp"
    src.write_text(TU)
    cmd = ["g++", "-std=c++11", "-fsyntax-only", "-w", "-fpermissive",
           "-I" + os.path.join(CFG, "Parallel"), "-I" + os.path.join(CFG, "a"),
           "-I" + os.path.join(CFG, "a", "b"), "-I" + os.path.join(CFG, "a", "b", "c"),
           "-I" + REF, "-I" + os.path.join(ROOT, "include"), str(src)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
