"""The K-cycle (psb_stop.ac_kcycle: two flexible-CG steps as the coarse
correction of chosen levels) against the oracle's restatement of the same
preconditioner: iteration counts +-1 and the solution to 1e-6 (the iterates
differ only by the order of the dot-product sums) in FP64, and in the mixed
solve the FP64 answer to the tolerance."""
import numpy as np
import pytest

import psb
import psb_oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def demeaned_rel(a, b):
    return rel(a - a.mean(), b - b.mean())


def system(n, nsph=10, seed=5):
    d = psb.generate(n, n, n, nsph, 0.05, 0.15, 1.0, 1000.0, seed, 1000 + seed)
    s = psb.Solver(n, n, n)
    s.set_matrix(d)
    return d, s


AC = dict(precond=psb.PSB_PREC_AC, norm=psb.PSB_NORM_L2REL, res_rat=1e-8, res_tol=0.0,
          max_iter=300)


@pytest.mark.parametrize("mask,flex", [(0b10, -1), (0b110, -1), (0b110, 0), (0, 1)])
def test_kcycle_fp64_matches_oracle(mask, flex):
    """flex: the outer CG's beta, -1 the default (flexible with the K-cycle,
    Fletcher-Reeves with the V-cycle), 0 Fletcher-Reeves, 1 flexible."""
    n = 64
    print(f"[kcycle] fp64 {n}^3 mask {mask:b} flexible {flex}", flush=True)
    d, s = system(n)
    b = d["rhs"]
    x, r = s.solve(np.zeros(b.size), b, ac_kcycle=mask, ac_flexible=flex, **AC)
    A = O.Mat(n, n, n, {k: d[k] for k in "cwesnbt"})
    ac = O.AC(A, np.ones(b.size), rule=1)
    try:
        O.set_kcycle(mask)
        O.set_flexible(flex == 1 or (flex == -1 and mask != 0))
        xo, info = ac.acpcg(np.zeros(b.size), b, rtol=1e-8, maxit=300, nu1=2, nu2=2, omega=0.9,
                            cfac=2.0)
    finally:
        O.set_kcycle(0)
        O.set_flexible(False)
    assert r["converged"] and info["converged"]
    assert abs(r["iterations"] - info["iters"]) <= 1, (r["iterations"], info["iters"])
    assert r["rel_l2"] < 1.2e-8
    assert demeaned_rel(x, xo) < 1e-6
    h = s.history()
    m = min(len(h), len(info["hist"]), 8)
    np.testing.assert_allclose(h[:m] / h[0], info["hist"][:m], rtol=1e-6)


@pytest.mark.parametrize("adapt", [2, 0])
def test_kcycle_mixed_reaches_the_fp64_answer(adapt):
    """With reliable updates (adapt 2) and with restarted refinement (adapt 0,
    inner solves to 1e-4)."""
    n = 96
    print(f"[kcycle] mixed {n}^3 adapt {adapt}", flush=True)
    d, s = system(n, nsph=16, seed=2)
    b = d["rhs"]
    x64, r64 = s.solve(np.zeros(b.size), b, precision=psb.PSB_FP64,
                       **dict(AC, res_rat=1e-12))
    xd, rd = s.solve(np.zeros(b.size), b, precision=psb.PSB_FP64, ac_kcycle=0b110, **AC)
    xk, rk = s.solve(np.zeros(b.size), b, precision=psb.PSB_MIXED, inner_adapt=adapt,
                     inner_rtol=1e-2 if adapt == 2 else 1e-4, ac_store=1, ac_kcycle=0b110, **AC)
    assert rd["converged"] and rk["converged"]
    assert rk["rel_l2"] < 1.2e-8
    if adapt == 2:  # the FP32 iterations with FP16 smoother copies track the FP64 ones
        assert abs(rk["iterations"] - rd["iterations"]) <= 2, (rk["iterations"], rd["iterations"])
    ed, ek = demeaned_rel(xd, x64), demeaned_rel(xk, x64)
    assert ek < 3.0 * ed + 1e-9, (ed, ek)


def test_kcycle_levels_chosen_by_size():
    """ac_kcycle = -1 takes the levels of 16^3 to 64^3 cells between the finest
    and the coarsest: at 128^3 (levels 128, 64, 32, 16, 8) levels 1-3."""
    n = 128
    print(f"[kcycle] auto {n}^3", flush=True)
    d, s = system(n, nsph=12, seed=3)
    b = d["rhs"]
    xa, ra = s.solve(np.zeros(b.size), b, ac_kcycle=-1, **AC)
    xm, rm = s.solve(np.zeros(b.size), b, ac_kcycle=0b1110, **AC)
    assert np.array_equal(xa, xm) and ra == rm


def test_kcycle_levels_of_semi_coarsened_hierarchy():
    """ac_kcycle = -1 with ac_coarsen = 1 takes the first four coarse levels."""
    nx, ny, nz = 64, 32, 96  # levels 64x32x96, 32x16x96, 16x8x96, 8x8x96, 8x8x48, ...
    d = psb.generate(nx, ny, nz, 8, 0.02, 0.06, 1.0, 1000.0, 19, 1019, stretch_z=2.5, lz=2.0)
    s = psb.Solver(nx, ny, nz)
    s.set_matrix(d)
    b = d["rhs"]
    kw = dict(AC, ac_coarsen=1, ac_cfac=1.0, max_iter=400)
    xa, ra = s.solve(np.zeros(b.size), b, ac_kcycle=-1, **kw)
    xm, rm = s.solve(np.zeros(b.size), b, ac_kcycle=0b11110, **kw)
    assert ra["converged"] and np.array_equal(xa, xm) and ra == rm


def test_kcycle_arguments():
    d, s = system(16, nsph=2)
    b = d["rhs"]
    for bad in (1, 0b11, -2):
        with pytest.raises(psb.PsbError, match="EINVAL"):
            s.solve(np.zeros(b.size), b, ac_kcycle=bad, **AC)
    # bits beyond the hierarchy are ignored: the V-cycle's result
    x0, r0 = s.solve(np.zeros(b.size), b, **AC)
    x1, r1 = s.solve(np.zeros(b.size), b, ac_kcycle=1 << 20, **AC)
    assert np.array_equal(x0, x1) and r0 == r1


def test_kcycle_on_semi_coarsened_levels_matches_oracle():
    """The K-cycle on the hierarchy that halves x and y before z
    (ac_coarsen = 1, the oracle's rule 3) of a z-stretched grid with solid
    cells: FP64 iteration counts +-1 and solutions against the oracle."""
    nx, ny, nz = 64, 32, 96
    print("[kcycle] semi-coarsened", flush=True)
    d = psb.generate(nx, ny, nz, 8, 0.02, 0.06, 1.0, 1000.0, 17, 1017, stretch_z=2.5, lz=2.0,
                     cylinder=True, cyl_d=0.2)
    A = O.Mat(nx, ny, nz, {k: d[k] for k in "cwesnbt"})
    b, n = d["rhs"], d["rhs"].size
    s = psb.Solver(nx, ny, nz)
    if d["off"].any():
        s.set_active(d["off"])
    s.set_matrix(d)
    mask = 0b110
    x, r = s.solve(np.zeros(n), b, ac_coarsen=1, ac_cfac=1.0, ac_kcycle=mask,
                   **dict(AC, max_iter=400))
    off = d["off"] if d["off"].any() else None
    ac = O.AC(A, np.ones(n), off=off, rule=3)
    try:
        O.set_kcycle(mask)
        O.set_flexible(True)  # the default beta with the K-cycle
        xo, info = ac.acpcg(np.zeros(n), b, rtol=1e-8, maxit=400, nu1=2, nu2=2, omega=0.9,
                            cfac=1.0)
    finally:
        O.set_kcycle(0)
        O.set_flexible(False)
    assert r["converged"] and info["converged"]
    assert abs(r["iterations"] - info["iters"]) <= 1, (r["iterations"], info["iters"])
    fl = d["off"] == 0
    assert demeaned_rel(x[fl], xo[fl]) < 1e-6


def test_mixed_kcycle_stretched_cylinder_against_oracle_fp64():
    """The stretched-grid workload's method in small: a z-stretched grid
    (tanh clustering 2.5) with a solid cylinder and bubbles, levels halving x
    and y before z, the K-cycle on the first four coarse levels, FP32
    iterations with FP16 smoother copies and FP64 reliable updates, to 1e-9.
    Against the oracle's FP64 solve of the same preconditioned method to
    1e-12: the GPU's FP64 residual meets the tolerance, the iteration count is
    the oracle's to 1e-9 within 10 %, and the solution (fluid cells, mean
    removed) is within 10 times the error the oracle's own FP64 solve to 1e-9
    leaves."""
    nx, ny, nz = 48, 32, 96
    print("[kcycle] mixed stretched cylinder", flush=True)
    d = psb.generate(nx, ny, nz, 12, 0.02, 0.06, 1.0, 1000.0, 23, 1023, stretch_z=2.5, lz=2.0,
                     cylinder=True, cyl_d=0.2)
    assert d["off"].any()
    A = O.Mat(nx, ny, nz, {k: d[k] for k in "cwesnbt"})
    b, n = d["rhs"], d["rhs"].size
    s = psb.Solver(nx, ny, nz)
    s.set_active(d["off"])
    s.set_matrix(d)
    x, r = s.solve(np.zeros(n), b, precision=psb.PSB_MIXED, inner_adapt=2, inner_rtol=1e-2,
                   ac_store=1, ac_coarsen=1, ac_cfac=1.0, ac_kcycle=-1,
                   **dict(AC, res_rat=1e-9, max_iter=600))
    ac = O.AC(A, np.ones(n), off=d["off"], rule=3)
    kw = dict(nu1=2, nu2=2, omega=0.9, cfac=1.0, maxit=600)
    try:
        O.set_kcycle(0b11110)  # the first four coarse levels, as ac_kcycle = -1 takes them
        O.set_flexible(True)
        x12, i12 = ac.acpcg(np.zeros(n), b, rtol=1e-12, **kw)
        x9, i9 = ac.acpcg(np.zeros(n), b, rtol=1e-9, **kw)
    finally:
        O.set_kcycle(0)
        O.set_flexible(False)
    assert r["converged"] and i12["converged"] and i9["converged"]
    assert r["rel_l2"] <= 1.05e-9, r
    assert abs(r["iterations"] - i9["iters"]) <= max(2, 0.1 * i9["iters"]), (r, i9["iters"])
    fl = d["off"] == 0
    e_gpu = demeaned_rel(x[fl], x12[fl])
    e_ora = demeaned_rel(x9[fl], x12[fl])
    print(f"[kcycle] error gpu mixed {e_gpu:.2e}, oracle fp64 to 1e-9 {e_ora:.2e}", flush=True)
    assert e_gpu <= 10.0 * e_ora + 1e-12, (e_gpu, e_ora)
