"""The marching stencil kernels (a block walks a slab of i planes of one
(j, k) tile) on shapes whose planes do not fill whole tiles, whose slabs do
not divide nx, and with few, large planes -- all above the size where the
library switches them on (2^15 cells).  The SpMV must stay bit-exact with the
oracle (the reference's four-pass order), CG with the diagonal
preconditioner must follow the oracle's iterates, and the AC-preconditioned
solves (FP64, and FP32 with reliable updates, whose Jacobi sweeps, CG p/q
steps and restrictions march) must converge to the tolerance."""
import numpy as np
import pytest

import psb
import psb_oracle as O
from test_gpu_callers import staggered_system

pytestmark = pytest.mark.gpu

SHAPES = [(40, 36, 44), (33, 70, 20), (6, 200, 64), (130, 18, 16)]


@pytest.mark.parametrize("shape", SHAPES)
def test_marching_spmv_and_cg_match_oracle(shape):
    nx, ny, nz = shape
    n = nx * ny * nz
    assert n >= 1 << 15
    A = staggered_system(nx, ny, nz, 11)
    s = psb.Solver(nx, ny, nz)
    s.set_matrix(A)
    Ao = O.Mat(nx, ny, nz, A)
    x = np.random.default_rng(12).standard_normal(n)
    assert np.array_equal(s.apply(x), O.spmv(Ao, x))
    b = np.random.default_rng(13).standard_normal(n)
    dV = np.ones(n)
    x4o, _ = O.cg(Ao, np.zeros(n), b, dV, maxit=4, res_rat=1e-14, res_tol=0.0)
    x4, _ = s.solve(np.zeros(n), b, max_iter=4, res_rat=1e-14, res_tol=0.0)
    assert np.linalg.norm(x4 - x4o) <= 1e-13 * np.linalg.norm(x4o)
    xo, io = O.cg(Ao, np.zeros(n), b, dV, maxit=500, res_rat=1e-10, res_tol=0.0)
    xg, r = s.solve(np.zeros(n), b, max_iter=500, res_rat=1e-10, res_tol=0.0)
    assert r["converged"] and abs(r["iterations"] - io["iters"]) <= 1


@pytest.mark.parametrize("shape", [(40, 36, 44), (64, 48, 32)])
def test_marching_ac_solves_converge(shape):
    nx, ny, nz = shape
    d = psb.generate(nx, ny, nz, 5, 0.05, 0.15, 1.0, 1000.0, 31, 1031)
    s = psb.Solver(nx, ny, nz)
    s.set_matrix(d)
    s.set_geometry(d["dxc"], d["dyc"], d["dzc"])
    b = d["rhs"]
    kw = dict(precond=psb.PSB_PREC_AC, norm=psb.PSB_NORM_L2REL, res_rat=1e-9, res_tol=0.0,
              max_iter=300)
    x64, r64 = s.solve(np.zeros(b.size), b, precision=psb.PSB_FP64, **kw)
    xm, rm = s.solve(np.zeros(b.size), b, precision=psb.PSB_MIXED, inner_adapt=2,
                     inner_rtol=1e-2, ac_store=1, **kw)
    for x, r in ((x64, r64), (xm, rm)):
        assert r["converged"]
        assert np.linalg.norm(s.apply(x, b)) <= 1e-9 * np.linalg.norm(b) * 1.0001
    assert abs(rm["iterations"] - r64["iterations"]) <= 3
