"""The LDS-staged marching kernels (psb_lds.h: the fine levels' Jacobi
sweeps, including the first post-sweep with the prolongation and the last
with the rho partials, and the CG p/q step) give the register-marching
kernels' results bit for bit: the same solves on a context created with
PSB_NO_LDS set (the register form) and on one without, on shapes whose rows
tile a block in one, several or a partial block of rows."""
import os

import numpy as np
import pytest

import psb

pytestmark = pytest.mark.gpu

SHAPES = [(64, 48, 32), (6, 200, 64), (130, 18, 16), (72, 40, 128), (40, 20, 512)]
AC = dict(precond=psb.PSB_PREC_AC, norm=psb.PSB_NORM_L2REL, res_rat=1e-9, res_tol=0.0,
          max_iter=300)
RUNS = [
    ("cg-di fp64", dict(max_iter=3000, res_rat=1e-8, res_tol=0.0)),
    ("ac fp64 kcycle", dict(AC, precision=psb.PSB_FP64, ac_kcycle=-1)),
    ("ac mixed fp16 smoother kcycle", dict(AC, precision=psb.PSB_MIXED, inner_adapt=2,
                                           inner_rtol=1e-2, ac_store=1, ac_kcycle=-1)),
    ("ac mixed fp32 smoother vcycle", dict(AC, precision=psb.PSB_MIXED, inner_adapt=2,
                                           inner_rtol=1e-2, ac_store=0)),
]


def run(d, shape, lds, kw):
    if lds:
        os.environ.pop("PSB_NO_LDS", None)
    else:
        os.environ["PSB_NO_LDS"] = "1"
    try:
        s = psb.Solver(*shape)
    finally:
        os.environ.pop("PSB_NO_LDS", None)
    s.set_matrix(d)
    s.set_geometry(d["dxc"], d["dyc"], d["dzc"])
    db = psb.DeviceArray.from_host(d["rhs"])
    dx = psb.DeviceArray(int(np.prod(shape)))
    dx.zero()
    r = s.solve_device(dx.addr, db.addr, **kw)
    x = dx.download()
    s.close()
    return x, r


@pytest.mark.parametrize("shape", SHAPES)
def test_lds_kernels_match_register_march(shape):
    d = psb.generate(*shape, 6, 0.05, 0.15, 1.0, 1000.0, 41, 1041)
    for name, kw in RUNS:
        print(f"[lds] {shape} {name}", flush=True)
        x1, r1 = run(d, shape, True, kw)
        x0, r0 = run(d, shape, False, kw)
        assert r1["iterations"] == r0["iterations"] and r1["outer"] == r0["outer"], (name, r1, r0)
        assert np.array_equal(x1, x0), (name, np.abs(x1 - x0).max())
        assert r1["converged"], (name, r1)
