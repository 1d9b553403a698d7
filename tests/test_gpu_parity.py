"""Parity of the HIP path (libpsb.so, through its C ABI) with the reference's
own outputs (golden fixtures) and with the oracle.

Tolerances: SpMV / residual are bit-exact (same per-cell rounding as the
reference's Scalar operators).  Krylov iterates differ only through the
order of the dot-product sums (tree vs sequential), so iterates after a few
iterations agree to 1e-12 relative, residual histories to 1e-9 relative and
iteration counts to +-1."""
import numpy as np
import pytest

import psb
import psb_oracle as O
from conftest import cycle_matrix, load_cycles, load_golden
from test_oracle_acbicgs import nonsym

pytestmark = pytest.mark.gpu


def solver_for(g, ghost=0):
    s = psb.Solver(*g.shape, ghost=ghost)
    s.set_matrix(g.mat())
    s.set_geometry(g["dxc"], g["dyc"], g["dzc"])
    return s


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def test_spmv_and_residual_bit_exact_vs_reference(golden):
    s = solver_for(golden)
    assert np.array_equal(s.apply(golden["x_rand"]), golden["Ax_rand"])
    assert np.array_equal(s.apply(golden["x_rand"], golden["b"]), golden["r_rand"])


def test_cg_diag_matches_reference(golden):
    s = solver_for(golden)
    b = golden["b"]
    n = b.size
    for k in golden.meta["cg_snapshots"]:
        x, r = s.solve(np.zeros(n), b, max_iter=k, res_rat=1e-8)
        assert r["iterations"] == k + 1 and not r["converged"]
        assert rel(x, golden["cg_x%d" % k]) < 1e-12
    x, r = s.solve(np.zeros(n), b, max_iter=2000, res_rat=1e-8)
    res, ratio, iters = golden.info["cg"][-1]
    assert r["converged"] and r["status"] == 0
    assert abs(r["iterations"] - iters) <= 1
    assert rel(x, golden["cg_xfull"]) < 1e-7
    # residual history against the oracle's (same norm, same stopping rule)
    A = O.Mat(*golden.shape, golden.mat())
    xo, info = O.cg(A, np.zeros(n), b, golden["dV"], maxit=2000, res_rat=1e-8)
    h = s.history()
    m = min(len(h), len(info["hist"]), 30)
    np.testing.assert_allclose(h[:m], info["hist"][:m], rtol=1e-9)
    assert r["res0"] == pytest.approx(info["res0"], rel=1e-13)


def test_bicgs_diag_matches_reference(golden):
    """BiCGS iterates match the reference's in the first iterations; the full
    iteration count is rounding-chaotic on these density-ratio-1000 systems
    (on bubbles16 the same algorithm needs 173 iterations with the
    reference's sequential sums, 172 with exact sums and 207 with numpy's
    pairwise sums), so it is bounded to 25 % and the converged solution to
    the accuracy the stopping rule gives."""
    s = solver_for(golden)
    b = golden["b"]
    n = b.size
    x, r = s.solve(np.zeros(n), b, max_iter=1, res_rat=1e-8, method=psb.PSB_BICGS)
    assert rel(x, golden["bicgs_x1"]) < 1e-12
    A = O.Mat(*golden.shape, golden.mat())
    x5o, _ = O.bicgs(A, np.zeros(n), b, golden["dV"], maxit=5, res_rat=1e-8)
    x, r = s.solve(np.zeros(n), b, max_iter=5, res_rat=1e-8, method=psb.PSB_BICGS)
    assert rel(x, x5o) < 1e-9
    x, r = s.solve(np.zeros(n), b, max_iter=2000, res_rat=1e-8, method=psb.PSB_BICGS)
    res, ratio, iters = golden.info["bicgs"][-1]
    assert r["converged"] and r["res"] < r["res0"] * 1e-8
    assert abs(r["iterations"] - iters) <= max(2, 0.25 * iters)
    xr = golden["bicgs_xfull"]
    assert rel(x - x.mean(), xr - xr.mean()) < 1e-5


def test_reference_ghost_layout():
    g = load_golden("bubbles16x12x8")
    nx, ny, nz = g.shape
    G = 3

    def pad(a):
        out = np.full((nx + 2 * G, ny + 2 * G, nz + 2 * G), 7.0)
        out[G:-G, G:-G, G:-G] = a.reshape(nx, ny, nz)
        return out.ravel()

    s = psb.Solver(nx, ny, nz, ghost=G)
    s.set_matrix({k: pad(g["A_" + k]) for k in "cwesnbt"})
    s.set_geometry(g["dxc"], g["dyc"], g["dzc"])
    y = s.apply(pad(g["x_rand"])).reshape(nx + 2 * G, ny + 2 * G, nz + 2 * G)
    assert np.array_equal(y[G:-G, G:-G, G:-G].ravel(), g["Ax_rand"])
    x, r = s.solve(pad(np.zeros(nx * ny * nz)), pad(g["b"]), max_iter=5, res_rat=1e-8)
    x = x.reshape(nx + 2 * G, ny + 2 * G, nz + 2 * G)
    assert rel(x[G:-G, G:-G, G:-G].ravel(), g["cg_x5"]) < 1e-12
    assert np.all(x[:G] == 7.0)  # ghosts untouched


def test_nonsymmetric_seven_array_path_and_odd_sizes():
    nx, ny, nz = 9, 10, 7
    rng = np.random.default_rng(3)
    n = nx * ny * nz
    A = {k: rng.uniform(0.1, 1.0, n) for k in "wesnbt"}
    g = lambda a: a.reshape(nx, ny, nz)
    g(A["w"])[0] = 0; g(A["e"])[-1] = 0; g(A["s"])[:, 0] = 0; g(A["n"])[:, -1] = 0
    g(A["b"])[:, :, 0] = 0; g(A["t"])[:, :, -1] = 0
    A["c"] = sum(A[k] for k in "wesnbt") + 0.5
    s = psb.Solver(nx, ny, nz)
    s.set_matrix(A)
    Am = O.Mat(nx, ny, nz, A)
    x = rng.standard_normal(n)
    b = rng.standard_normal(n)
    assert np.array_equal(s.apply(x), O.spmv(Am, x))
    assert np.array_equal(s.apply(x, b), O.resid(Am, b, x))
    dV = np.ones(n)
    xg, r = s.solve(np.zeros(n), b, max_iter=200, res_rat=1e-10, method=psb.PSB_BICGS)
    xo, info = O.bicgs(Am, np.zeros(n), b, dV, maxit=200, res_rat=1e-10)
    assert r["converged"] == info["converged"]
    assert abs(r["iterations"] - info["iters"]) <= 2
    assert rel(xg, xo) < 1e-8


def test_stopping_rules_match_oracle():
    g = load_golden("bubbles16")
    s = solver_for(g)
    A = O.Mat(*g.shape, g.mat())
    b, dV, n = g["b"], g["dV"], g["b"].size
    # MaxIter exhausted: converged false, printed count MaxIter+1 (cg.cpp:178)
    x, r = s.solve(np.zeros(n), b, max_iter=7, res_rat=1e-12)
    assert r["status"] == 1
    assert r["iterations"] == 8 and not r["converged"]
    # ResTol above the initial residual: immediate true (cg.cpp:64)
    x, r = s.solve(np.zeros(n), b, max_iter=50, res_tol=1e6)
    assert r["converged"] and r["iterations"] == 0 and np.all(x == 0)
    # MinIter gates the ratio test
    xo, info = O.cg(A, np.zeros(n), b, dV, mini=30, maxit=200, res_rat=0.5)
    x, r = s.solve(np.zeros(n), b, min_iter=30, max_iter=200, res_rat=0.5)
    assert r["iterations"] == info["iters"] == 30
    # scale divides the norm
    x, r = s.solve(np.zeros(n), b, max_iter=3, scale=4.0)
    xo, info = O.cg(A, np.zeros(n), b, dV, maxit=3, scale=4.0)
    assert r["res0"] == pytest.approx(info["res0"], rel=1e-13)
    # staleness window
    xo, info = O.cg(A, np.zeros(n), b, dV, maxit=400, res_rat=1e-30, res_tol=0.0, stalecount=3)
    x, r = s.solve(np.zeros(n), b, max_iter=400, res_rat=1e-30, res_tol=0.0, stalecount=3)
    assert info["reason"] == 3 and r["status"] == 2
    assert abs(r["iterations"] - info["iters"]) <= 3


def test_nan_input_is_reported():
    g = load_golden("drop12")
    s = solver_for(g)
    b = g["b"].copy()
    b[5] = np.nan
    x, r = s.solve(np.zeros(b.size), b, max_iter=10)
    assert r["status"] == 5 and not r["converged"]


def ac_case(n, nsph=4, ratio=1000.0, seed=2, ny=None, nz=None):
    ny, nz = ny or n, nz or n
    d = psb.generate(n, ny, nz, nsph, 0.1, 0.25, 1.0, ratio, seed, 1000 + seed)
    A = O.Mat(n, ny, nz, {k: d[k] for k in "cwesnbt"})
    return d, A


@pytest.mark.parametrize("shape", [(32, 32, 32), (48, 32, 16)])
def test_acpcg_fp64_matches_oracle(shape):
    nx, ny, nz = shape
    d, A = ac_case(nx, ny=ny, nz=nz)
    b = d["rhs"]
    n = b.size
    s = psb.Solver(nx, ny, nz)
    s.set_matrix(d)
    x, r = s.solve(np.zeros(n), b, precond=psb.PSB_PREC_AC, norm=psb.PSB_NORM_L2REL,
                   res_rat=1e-8, max_iter=200)
    ac = O.AC(A, np.ones(n), rule=1)
    xo, info = ac.acpcg(np.zeros(n), b, rtol=1e-8, maxit=200, nu1=2, nu2=2, omega=0.9, cfac=2.0)
    assert r["converged"] and info["converged"]
    assert abs(r["iterations"] - info["iters"]) <= 1
    assert r["rel_l2"] < 1.2e-8
    assert rel(x - x.mean(), xo - xo.mean()) < 1e-6


@pytest.mark.parametrize("case", ["bubbles16", "gen32", "nonsym16", "nonsym32"])
def test_ac_bicgs_fp64_matches_oracle(case):
    """BiCGS with the V-cycle as its preconditioner (FP64, method PSB_BICGS
    with PSB_PREC_AC) against the oracle's restatement: the first iterates
    and residual histories to the order of the dot-product sums, the
    converged solve within an iteration and to the stopping rule's
    accuracy.  nonsym*: non-symmetric matrices (7-array levels, the
    coarsest inverse by LU): the reference driver's matrix of its
    non-symmetric cycle runs, and the same pattern on a 32^3 case."""
    if case in ("bubbles16", "nonsym16"):
        g = load_golden("bubbles16")
        M = g.mat() if case == "bubbles16" else \
            cycle_matrix(g, "cycles_nonsym", load_cycles("bubbles16", "cycles_nonsym")[0])
        s = psb.Solver(*g.shape)
        s.set_matrix(M)
        s.set_geometry(g["dxc"], g["dyc"], g["dzc"])
        A = O.Mat(*g.shape, M)
        b, dV = g["b"], g["dV"]
    else:
        d, A = ac_case(32, nsph=5, seed=7)
        b, dV = d["rhs"], np.ones(d["rhs"].size)
        M = d
        if case == "nonsym32":
            A = nonsym(A, (32, 32, 32))
            M = A.a
        s = psb.Solver(32, 32, 32)
        s.set_matrix(M)
    n = b.size
    ac = O.AC(A, dV, rule=1)
    kw = dict(precond=psb.PSB_PREC_AC, method=psb.PSB_BICGS, res_tol=0.0)
    okw = dict(nu1=2, nu2=2, omega=0.9, cfac=2.0, res_tol=0.0)
    for k in (1, 3):
        x, r = s.solve(np.zeros(n), b, max_iter=k, res_rat=1e-14, **kw)
        xo, io = ac.acbicgs(np.zeros(n), b, dV, maxit=k, res_rat=1e-14, **okw)
        assert r["iterations"] == io["iters"] == k + 1
        assert rel(x, xo) < 1e-10, (k, rel(x, xo))
        assert np.allclose(s.history(), io["hist"], rtol=1e-9, atol=0)
    x, r = s.solve(np.zeros(n), b, max_iter=300, res_rat=1e-10, **kw)
    xo, io = ac.acbicgs(np.zeros(n), b, dV, maxit=300, res_rat=1e-10, **okw)
    assert r["converged"] and io["converged"] and r["status"] == psb.PSB_OK
    assert abs(r["iterations"] - io["iters"]) <= 1, (r["iterations"], io["iters"])
    assert r["iterations"] < 40
    if case.startswith("nonsym"):  # non-singular: x itself
        assert rel(x, xo) < 1e-7
    else:
        assert rel(x - x.mean(), xo - xo.mean()) < 1e-7
    h = s.history()
    assert len(h) == r["iterations"] + 1 and h[-1] < 1e-10 * h[0]
    # the FP32 cycle only for symmetric matrices
    if case.startswith("nonsym"):
        with pytest.raises(RuntimeError):
            s.solve(np.zeros(n), b, precision=psb.PSB_MIXED, **kw)


@pytest.mark.parametrize("sym", [True, False])
def test_ac_bicgs_semi_coarsened_with_solid_cells_matches_oracle(sym):
    """The V-cycle-preconditioned BiCGS on the hierarchy that halves x and y
    before z (ac_coarsen = 1, the oracle's rule 3) of a z-stretched grid with
    a solid cylinder, on the pressure matrix and on it made non-symmetric:
    iteration counts +-1 and the solution on the fluid cells against the
    oracle; solid cells stay 0."""
    nx, ny, nz = 48, 32, 64
    print(f"[acbicgs] semi-coarsened sym={sym}", flush=True)
    d = psb.generate(nx, ny, nz, 8, 0.02, 0.06, 1.0, 1000.0, 17, 1017, stretch_z=2.5, lz=2.0,
                     cylinder=True, cyl_d=0.2)
    A = O.Mat(nx, ny, nz, {k: d[k] for k in "cwesnbt"})
    if not sym:
        A = nonsym(A, (nx, ny, nz))
    b, n, off = d["rhs"], d["rhs"].size, d["off"]
    assert off.any()
    s = psb.Solver(nx, ny, nz)
    s.set_active(off)
    s.set_matrix(A.a)
    x, r = s.solve(np.zeros(n), b, precond=psb.PSB_PREC_AC, method=psb.PSB_BICGS, ac_coarsen=1,
                   ac_cfac=1.0, norm=psb.PSB_NORM_L2REL, res_rat=1e-8, res_tol=0.0,
                   max_iter=300)
    ac = O.AC(A, np.ones(n), off=off, rule=3)
    xo, io = ac.acbicgs(np.zeros(n), b, np.ones(n), maxit=300, res_rat=1e-8, res_tol=0.0,
                        nu1=2, nu2=2, omega=0.9, cfac=1.0)
    assert r["converged"] and io["converged"]
    assert abs(r["iterations"] - io["iters"]) <= 1, (r["iterations"], io["iters"])
    fl = off == 0
    if sym:
        assert rel(x[fl] - x[fl].mean(), xo[fl] - xo[fl].mean()) < 1e-6
    else:
        assert rel(x[fl], xo[fl]) < 1e-6
    assert np.all(x[off == 1] == 0.0)


@pytest.mark.parametrize("case", ["bubbles16", "gen48"])
def test_ac_bicgs_with_fp32_cycle_matches_fp64(case):
    """The preconditioned BiCGS with its V-cycle on the FP32 levels
    (precision PSB_MIXED): the BiCGS recurrences stay FP64, so the true
    residual meets the tolerance and the solution is the FP64 form's to it;
    the FP32 preconditioner costs at most a few iterations."""
    if case == "bubbles16":
        g = load_golden(case)
        s = solver_for(g)
        b = g["b"]
    else:
        d, _ = ac_case(48, nsph=6, seed=3)
        s = psb.Solver(48, 48, 48)
        s.set_matrix(d)
        b = d["rhs"]
    n = b.size
    kw = dict(precond=psb.PSB_PREC_AC, method=psb.PSB_BICGS, norm=psb.PSB_NORM_L2REL,
              res_rat=1e-10, res_tol=0.0, max_iter=300)
    x64, r64 = s.solve(np.zeros(n), b, **kw)
    x32, r32 = s.solve(np.zeros(n), b, precision=psb.PSB_MIXED, **kw)
    assert r64["converged"] and r32["converged"] and r32["status"] == psb.PSB_OK
    assert r32["iterations"] <= r64["iterations"] + max(2, r64["iterations"] // 5), \
        (r32["iterations"], r64["iterations"])
    assert r32["rel_l2"] <= 1.05e-10
    assert rel(x32 - x32.mean(), x64 - x64.mean()) < 1e-8
    # FP64 again on the same context: its own graph, the same result
    x2, r2 = s.solve(np.zeros(n), b, **kw)
    assert r2["iterations"] == r64["iterations"] and np.array_equal(x2, x64)


def test_ac_bicgs_singular_nonsymmetric_matrix_is_rejected():
    """A non-symmetric Neumann matrix without a time term is singular on its
    coarsest level: PSB_EINVAL, as the oracle reports it."""
    d, A = ac_case(32, nsph=5, seed=7)
    M = {q: np.array(d[q], dtype=np.float64) for q in "cwesnbt"}
    i = np.meshgrid(np.arange(32), np.arange(32), np.arange(32), indexing="ij")[0].ravel()
    de = 0.2 * (i % 3) * M["e"]
    M["e"] = M["e"] + de
    M["c"] = M["c"] + de
    s = psb.Solver(32, 32, 32)
    s.set_matrix(M)
    with pytest.raises(RuntimeError, match="EINVAL|invalid"):
        s.solve(np.zeros(32 ** 3), d["rhs"], precond=psb.PSB_PREC_AC, method=psb.PSB_BICGS)
    with pytest.raises(ValueError):
        O.AC(O.Mat(32, 32, 32, M), np.ones(32 ** 3), rule=1).coarse_dense()
    # the context stays usable: BiCGS-di on the same matrix
    x, r = s.solve(np.zeros(32 ** 3), d["rhs"], method=psb.PSB_BICGS, max_iter=3)
    assert r["iterations"] == 4


def test_mixed_precision_refinement_reaches_fp64_accuracy():
    d, A = ac_case(48, nsph=6, seed=3)
    b = d["rhs"]
    n = b.size
    s = psb.Solver(48, 48, 48)
    s.set_matrix(d)
    x, r = s.solve(np.zeros(n), b, precond=psb.PSB_PREC_AC, precision=psb.PSB_MIXED,
                   norm=psb.PSB_NORM_L2REL, res_rat=1e-10, max_iter=500, inner_rtol=1e-4)
    assert r["converged"] and r["outer"] >= 2
    assert r["rel_l2"] <= 1e-10 * 1.01
    # against an FP64 solve to 1e-12 by the oracle
    ac = O.AC(A, np.ones(n), rule=1)
    xo, info = ac.acpcg(np.zeros(n), b, rtol=1e-12, maxit=400)
    err = rel(x - x.mean(), xo - xo.mean())
    assert err < 1e-7


def test_reliable_update_mixed_matches_fp64_iterations():
    """FP32 AC-PCG with reliable updates: the FP64 solution to the FP64
    accuracy, in as many iterations as the FP64 AC-PCG (+-2)."""
    d, A = ac_case(48, nsph=6, seed=3)
    b = d["rhs"]
    n = b.size
    s = psb.Solver(48, 48, 48)
    s.set_matrix(d)
    kw = dict(precond=psb.PSB_PREC_AC, norm=psb.PSB_NORM_L2REL, res_rat=1e-10, max_iter=500)
    _, r64 = s.solve(np.zeros(n), b, **kw)
    x, r = s.solve(np.zeros(n), b, precision=psb.PSB_MIXED, inner_adapt=2, inner_rtol=1e-2,
                   **kw)
    assert r["converged"] and r["status"] == psb.PSB_OK and r["outer"] >= 3
    assert abs(r["iterations"] - r64["iterations"]) <= 2
    assert r["rel_l2"] <= 1e-10 * 1.01
    # the reported residual is the true one of the returned x
    assert abs(r["res"] / r["res0"] - r["rel_l2"]) <= 1e-6 * r["rel_l2"]
    h = s.history()
    assert len(h) == r["outer"] + 1 and h[0] == r["res0"] and h[-1] == r["res"]
    # each update after a 1e-2 drop of the recursive residual, the last one
    # when it reaches the target
    assert np.all(h[1:-1] < h[:-2] * 0.02) and h[-1] < h[-2]
    ac = O.AC(A, np.ones(n), rule=1)
    xo, info = ac.acpcg(np.zeros(n), b, rtol=1e-12, maxit=400)
    assert rel(x - x.mean(), xo - xo.mean()) < 1e-7


def test_reliable_update_max_iter_reports_true_residual():
    d, _ = ac_case(32, nsph=4, seed=5)
    b = d["rhs"]
    n = b.size
    s = psb.Solver(32, 32, 32)
    s.set_matrix(d)
    x, r = s.solve(np.zeros(n), b, precond=psb.PSB_PREC_AC, precision=psb.PSB_MIXED,
                   inner_adapt=2, inner_rtol=1e-3, norm=psb.PSB_NORM_L2REL, res_rat=1e-12,
                   max_iter=5)
    assert not r["converged"] and r["status"] == psb.PSB_MAXITER
    assert r["iterations"] == 6  # reference count maxit + 1 (cg.cpp:178)
    assert abs(r["res"] / r["res0"] - r["rel_l2"]) <= 1e-6 * r["rel_l2"]


@pytest.mark.parametrize("precision", [psb.PSB_FP64, psb.PSB_MIXED])
def test_ac_with_reference_norm(precision):
    """PSB_NORM_REF with the AC preconditioner on a stretched grid: the
    reported residual is sqrt(sum (r/dV)^2)/N of the returned x."""
    nx, ny, nz = 32, 16, 48
    d = psb.generate(nx, ny, nz, 6, 0.02, 0.06, 1.0, 1000.0, 11, 1011, stretch_z=2.0, lz=1.5)
    A = O.Mat(nx, ny, nz, {k: d[k] for k in "cwesnbt"})
    b, n = d["rhs"], d["rhs"].size
    s = psb.Solver(nx, ny, nz)
    s.set_matrix(d)
    s.set_geometry(d["dxc"], d["dyc"], d["dzc"])
    x, r = s.solve(np.zeros(n), b, precond=psb.PSB_PREC_AC, precision=precision,
                   inner_adapt=2, inner_rtol=1e-2, norm=psb.PSB_NORM_REF, res_rat=1e-9,
                   res_tol=0.0, max_iter=500)
    assert r["converged"]
    dV = (d["dxc"][:, None, None] * d["dyc"][None, :, None] * d["dzc"][None, None, :]).ravel()
    rr = O.resid(A, b, x)
    ref = np.sqrt(np.sum((rr / dV) ** 2)) / n
    ref0 = np.sqrt(np.sum((b / dV) ** 2)) / n
    assert abs(r["res0"] - ref0) <= 1e-12 * ref0
    tol = 1e-9 if precision == psb.PSB_MIXED else 1e-4  # FP64: recursive residual
    assert abs(r["res"] - ref) <= tol * ref
    assert r["res"] < 1e-9 * r["res0"]


def test_semi_coarsened_hierarchy_matches_oracle():
    """ac_coarsen = 1 (x and y halved before z) on a z-stretched grid: FP64
    AC-PCG against the oracle on the same levels (its rule 3)."""
    nx, ny, nz = 32, 16, 48
    d = psb.generate(nx, ny, nz, 6, 0.02, 0.06, 1.0, 1000.0, 13, 1013, stretch_z=2.5, lz=2.0)
    A = O.Mat(nx, ny, nz, {k: d[k] for k in "cwesnbt"})
    b, n = d["rhs"], d["rhs"].size
    s = psb.Solver(nx, ny, nz)
    s.set_matrix(d)
    kw = dict(precond=psb.PSB_PREC_AC, norm=psb.PSB_NORM_L2REL, res_rat=1e-8, res_tol=0.0,
              max_iter=400, ac_cfac=1.0)
    x, r = s.solve(np.zeros(n), b, ac_coarsen=1, **kw)
    ac = O.AC(A, np.ones(n), rule=3)
    xo, info = ac.acpcg(np.zeros(n), b, rtol=1e-8, maxit=400, nu1=2, nu2=2, omega=0.9, cfac=1.0)
    assert r["converged"] and info["converged"]
    assert abs(r["iterations"] - info["iters"]) <= 1, (r["iterations"], info["iters"])
    assert rel(x - x.mean(), xo - xo.mean()) < 1e-6
    # switching back rebuilds the isotropic hierarchy (oracle rule 1)
    _, r0 = s.solve(np.zeros(n), b, ac_coarsen=0, **kw)
    _, i1 = O.AC(A, np.ones(n), rule=1).acpcg(np.zeros(n), b, rtol=1e-8, maxit=400, nu1=2,
                                              nu2=2, omega=0.9, cfac=1.0)
    assert abs(r0["iterations"] - i1["iters"]) <= 1, (r0["iterations"], i1["iters"])
    # and the mixed-precision solve on the semi-coarsened levels
    _, rm = s.solve(np.zeros(n), b, ac_coarsen=1, precision=psb.PSB_MIXED, inner_adapt=2,
                    inner_rtol=1e-2, **kw)
    assert rm["converged"] and rm["rel_l2"] <= 1e-8 * 1.01


@pytest.mark.parametrize("adapt,rtol", [(0, 1e-4), (2, 1e-2)])
def test_acpcg_with_solid_cells_and_stretched_grid(adapt, rtol):
    nx, ny, nz = 32, 16, 64
    d = psb.generate(nx, ny, nz, 8, 0.01, 0.04, 1.0, 1000.0, 7, 1007, stretch_z=2.5, lz=2.0,
                     cylinder=True, cyl_d=0.2)
    assert d["off"].sum() > 0
    b, n = d["rhs"], d["rhs"].size
    s = psb.Solver(nx, ny, nz)
    s.set_active(d["off"])
    s.set_matrix(d)
    s.set_geometry(d["dxc"], d["dyc"], d["dzc"])
    x, r = s.solve(np.zeros(n), b, precond=psb.PSB_PREC_AC, precision=psb.PSB_MIXED,
                   norm=psb.PSB_NORM_L2REL, res_rat=1e-9, max_iter=1000, inner_adapt=adapt,
                   inner_rtol=rtol)
    assert r["converged"] and r["rel_l2"] <= 1e-9 * 1.01
    off = d["off"].astype(bool)
    assert np.all(x[off] == 0.0)


def test_reference_ac_vcycle_matches_reference(golden):
    """The reference's AC::vcycle (CG smoothers on its own levels) on the
    GPU against the reference's own results: one cycle, and the full solve to
    ResRat 1e-8 (cycles, convergence, per-cycle residuals, solution).  Only
    the order of the dot-product sums differs (tree vs sequential), so
    iterates agree to rounding amplified by 20-100 CG iterations per level.
    The constant component is compared apart: 100 CG iterations on the
    coarsest (singular, Neumann) level run far past convergence and drift
    along the constants by an amount set by rounding alone (0.13 on a field
    of size 30 in the 16^3 case); the residual does not see it."""
    s = solver_for(golden)
    b = golden["b"]
    n = b.size
    x1, r1 = s.ac_vcycle(np.zeros(n), b, res_rat=1e-30, max_cycles=1)
    assert r1["iterations"] == 1
    ref1 = golden["ac_phi1"]
    assert rel(x1 - x1.mean(), ref1 - ref1.mean()) < 1e-8
    x, r = s.ac_vcycle(np.zeros(n), b, res_rat=1e-8, max_cycles=20)
    ref_full = [q for q in golden.info["acret"] if q[0] == "full"][0]
    assert r["iterations"] == ref_full[2] and bool(r["converged"]) == bool(ref_full[1])
    h = s.history()
    cyc = golden.info["cycles"][-ref_full[2]:]
    for c, res in cyc:
        assert h[c] == pytest.approx(res, rel=1e-7)
    reff = golden["ac_phifull"]
    assert rel(x - x.mean(), reff - reff.mean()) < 1e-7
    if r["converged"]:
        assert r["status"] == psb.PSB_OK and r["res"] < 1e-8 * r["res0"]


def test_reference_ac_vcycle_on_stretched_grid_with_solids_matches_oracle():
    nx, ny, nz = 24, 12, 18          # 18 = 2 * 3 * 3: levels by factors 2, 3
    d = psb.generate(nx, ny, nz, 4, 0.02, 0.06, 1.0, 100.0, 17, 1017, stretch_z=1.5, lz=1.5,
                     cylinder=True, cyl_d=0.4)
    assert d["off"].sum() > 0
    A = O.Mat(nx, ny, nz, {k: d[k] for k in "cwesnbt"})
    dV = (d["dxc"][:, None, None] * d["dyc"][None, :, None] * d["dzc"][None, None, :]).ravel()
    b, n = d["rhs"], d["rhs"].size
    s = psb.Solver(nx, ny, nz)
    s.set_active(d["off"])
    s.set_matrix(d)
    s.set_geometry(d["dxc"], d["dyc"], d["dzc"])
    ac = O.AC(A, dV, off=d["off"], rule=0)
    assert O.levels(0, nx, ny, nz) == [(24, 12, 18), (12, 6, 9), (6, 3, 3), (3, 3, 3)]
    xo, io = ac.vcycle(np.zeros(n), b, factor=1e-6, max_cyc=20, dt=0.5)
    x, r = s.ac_vcycle(np.zeros(n), b, res_rat=1e-6, max_cycles=20, dt=0.5)
    assert r["iterations"] == io["ncyc"] and bool(r["converged"]) == io["converged"]
    h = s.history()
    assert np.allclose(h[:len(io["hist"])], io["hist"], rtol=1e-6)
    fl = d["off"] == 0
    assert rel(x[fl] - x[fl].mean(), xo[fl] - xo[fl].mean()) < 1e-6


def test_cg_diag_64_drop_matches_oracle():
    """The plain reference method at its full size: 64^3 unit cube, one
    centred drop (radius 0.25, density 1000 in 1), CG with the diagonal
    preconditioner to 1e-8 in the reference norm.  Iteration count +-1 and
    solution within 1e-8 relative of the oracle's (the reference's algorithm)."""
    n = 64
    d = psb.generate(n, n, n, 1, 0.25, 0.25, 1000.0, 1.0, 0, 1001, centre=(0.5, 0.5, 0.5))
    A = O.Mat(n, n, n, {k: d[k] for k in "cwesnbt"})
    dV = np.full(n ** 3, (1.0 / n) ** 3)
    s = psb.Solver(n, n, n)
    s.set_matrix(d)
    s.set_geometry(d["dxc"], d["dyc"], d["dzc"])
    x, r = s.solve(np.zeros(n ** 3), d["rhs"], max_iter=1000, res_rat=1e-8)
    xo, info = O.cg(A, np.zeros(n ** 3), d["rhs"], dV, maxit=1000, res_rat=1e-8)
    assert r["converged"] and info["converged"]
    assert abs(r["iterations"] - info["iters"]) <= 1, (r["iterations"], info["iters"])
    assert r["res"] < 1e-8 * r["res0"]
    assert rel(x, xo) < 1e-8


@pytest.mark.parametrize("store", [0, 1])
def test_fused_post_smoothing_gives_the_same_solve(store):
    """The two post-sweeps in one pass (ac_fuse = 1) compute the same z as
    one kernel per sweep; only the block partition of the rho sums differs,
    so iteration counts agree and the solutions to rounding."""
    d, A = ac_case(64, nsph=6, seed=9)
    b = d["rhs"]
    n = b.size
    s = psb.Solver(64, 64, 64)
    s.set_matrix(d)
    kw = dict(precond=psb.PSB_PREC_AC, precision=psb.PSB_MIXED, inner_adapt=2, inner_rtol=1e-2,
              norm=psb.PSB_NORM_L2REL, res_rat=1e-9, res_tol=0.0, max_iter=300, ac_store=store)
    x1, r1 = s.solve(np.zeros(n), b, ac_fuse=1, **kw)
    x0, r0 = s.solve(np.zeros(n), b, ac_fuse=0, **kw)
    assert r1["converged"] and r0["converged"]
    assert abs(r1["iterations"] - r0["iterations"]) <= 1
    assert r1["rel_l2"] <= 1e-9 * 1.01
    assert rel(x1 - x1.mean(), x0 - x0.mean()) < 1e-7


@pytest.mark.parametrize("ac_store", [1, 2])
def test_fp16_level_copies_keep_the_iterations(ac_store):
    """FP16 copies of the level matrices for the smoother (ac_store = 1) and
    also for the FP32 CG operator (2): the reliable updates still measure the
    true FP64 residual, which meets the tolerance, in about the iterations of
    the FP32 copies (the same solve with ac_store = 0)."""
    n = 48
    d = psb.generate(n, n, n, 6, 0.05, 0.15, 1.0, 1000.0, 41, 1041)
    s = psb.Solver(n, n, n)
    s.set_matrix(d)
    s.set_geometry(d["dxc"], d["dyc"], d["dzc"])
    b = d["rhs"]
    kw = dict(precond=psb.PSB_PREC_AC, norm=psb.PSB_NORM_L2REL, res_rat=1e-9, res_tol=0.0,
              max_iter=300, precision=psb.PSB_MIXED, inner_adapt=2, inner_rtol=1e-2)
    x0, r0 = s.solve(np.zeros(b.size), b, ac_store=0, **kw)
    x1, r1 = s.solve(np.zeros(b.size), b, ac_store=ac_store, **kw)
    assert r0["converged"] and r1["converged"]
    assert np.linalg.norm(s.apply(x1, b)) <= 1e-9 * np.linalg.norm(b) * 1.0001
    assert abs(r1["iterations"] - r0["iterations"]) <= 3
    assert rel(x1 - x1.mean(), x0 - x0.mean()) < 1e-6
