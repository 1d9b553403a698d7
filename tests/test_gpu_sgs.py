"""CG with the red-black symmetric Gauss-Seidel preconditioner
(PSB_PREC_SGS), the library's parallel replacement for the reference's
Prec::ic0 / ic2, against the oracle's restatement of the same method (same
iterates up to the order of the dot-product sums: iterates after a few
iterations to 1e-11 relative, histories to 1e-9, iteration counts +-1), and
against the reference's own IC2 runs (<case>_prec.npz) where only the
answer is shared: the same solution to the 1e-8 tolerance, in 1.5-1.8 times
the iterations."""
import json
import os

import numpy as np
import pytest

import psb
import psb_oracle as O
from conftest import GOLDEN, PREC_CASES, load_golden

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def demeaned_rel(a, b):
    return rel(a - a.mean(), b - b.mean())


def solver_for(g):
    s = psb.Solver(*g.shape)
    s.set_matrix(g.mat())
    s.set_geometry(g["dxc"], g["dyc"], g["dzc"])
    return s


SGS = dict(precond=psb.PSB_PREC_SGS, res_tol=1e-15)


@pytest.mark.parametrize("case", PREC_CASES)
def test_sgs_cg_matches_oracle_and_reaches_the_ic2_answer(case):
    g = load_golden(case)
    print(f"[sgs] {case}", flush=True)
    z = np.load(os.path.join(GOLDEN, case + "_prec.npz"), allow_pickle=False)
    meta = json.loads(str(z["meta"]))
    s = solver_for(g)
    A = O.Mat(*g.shape, g.mat())
    b, n = g["b"], g["b"].size
    for k in (1, 5):
        x, r = s.solve(np.zeros(n), b, max_iter=k, res_rat=1e-8, **SGS)
        xo, info = O.cg_prec(A, "rbsgs", np.zeros(n), b, g["dV"], maxit=k, res_rat=1e-8,
                             res_tol=1e-15)
        assert r["iterations"] == info["iters"] == k + 1 and not r["converged"]
        assert rel(x, xo) < 1e-11
    x, r = s.solve(np.zeros(n), b, max_iter=2000, res_rat=1e-8, **SGS)
    xo, info = O.cg_prec(A, "rbsgs", np.zeros(n), b, g["dV"], maxit=2000, res_rat=1e-8,
                         res_tol=1e-15)
    assert r["converged"] and info["converged"] and r["status"] == psb.PSB_OK
    assert abs(r["iterations"] - info["iters"]) <= 1
    assert rel(x, xo) < 1e-7
    h = s.history()
    m = min(len(h), len(info["hist"]), 30)
    np.testing.assert_allclose(h[:m], info["hist"][:m], rtol=1e-9)
    # the reference's IC2: the same answer, fewer (sequential) iterations
    ic2_iters = meta["runs"]["ic2"][-1][2]
    assert ic2_iters < r["iterations"] <= 2 * ic2_iters
    assert demeaned_rel(x, z["ic2_xfull"]) < 1e-6


def test_sgs_cg_on_the_64_drop_and_nonsymmetric_storage():
    """64^3 unit cube with the centred drop (the reference's CG case at its
    size) through both the symmetric 4-array and the 7-array storage."""
    n = 64
    d = psb.generate(n, n, n, 1, 0.25, 0.25, 1000.0, 1.0, 0, 1001, centre=(0.5, 0.5, 0.5))
    A = O.Mat(n, n, n, {k: d[k] for k in "cwesnbt"})
    dV = np.full(n ** 3, (1.0 / n) ** 3)
    xo, info = O.cg_prec(A, "rbsgs", np.zeros(n ** 3), d["rhs"], dV, maxit=1000,
                         res_rat=1e-8, res_tol=1e-15)
    xd, idi = O.cg(A, np.zeros(n ** 3), d["rhs"], dV, maxit=1000, res_rat=1e-8,
                   res_tol=1e-15)
    assert info["converged"] and info["iters"] <= 0.6 * idi["iters"]
    for sym in (True, False):
        print(f"[sgs] 64 drop sym={sym}", flush=True)
        s = psb.Solver(n, n, n)
        M = {k: d[k] for k in "cwesnbt"}
        if not sym:  # break the exact symmetry by one ulp so the 7-array path runs
            M["e"] = M["e"].copy()
            M["e"][M["e"] > 0] *= 1.0 + 2.0 ** -52
        s.set_matrix(M)
        s.set_geometry(d["dxc"], d["dyc"], d["dzc"])
        x, r = s.solve(np.zeros(n ** 3), d["rhs"], max_iter=1000, res_rat=1e-8, **SGS)
        assert r["converged"]
        assert abs(r["iterations"] - info["iters"]) <= 1, (r["iterations"], info["iters"])
        assert rel(x, xo) < 1e-7


def test_sgs_cg_stopping_rules_and_arguments():
    g = load_golden("bubbles16")
    s = solver_for(g)
    A = O.Mat(*g.shape, g.mat())
    b, dV, n = g["b"], g["dV"], g["b"].size
    x, r = s.solve(np.zeros(n), b, max_iter=7, res_rat=1e-12, **SGS)
    assert r["status"] == psb.PSB_MAXITER and r["iterations"] == 8
    x, r = s.solve(np.zeros(n), b, max_iter=50, res_rat=0.01, res_tol=1e6,
                   precond=psb.PSB_PREC_SGS)
    assert r["converged"] and r["iterations"] == 0 and not x.any()
    xo, info = O.cg_prec(A, "rbsgs", np.zeros(n), b, dV, mini=20, maxit=200, res_rat=0.5,
                         res_tol=1e-15)
    x, r = s.solve(np.zeros(n), b, min_iter=20, max_iter=200, res_rat=0.5, **SGS)
    assert r["iterations"] == info["iters"] == 20
    with pytest.raises(psb.PsbError, match="EINVAL"):
        s.solve(np.zeros(n), b, method=psb.PSB_BICGS, **SGS)
    with pytest.raises(psb.PsbError, match="EINVAL"):
        s.solve(np.zeros(n), b, precision=psb.PSB_MIXED, **SGS)
    with pytest.raises(psb.PsbError, match="EINVAL"):
        s.solve(np.zeros(n), b, precond=7)
