"""The oracle's CG preconditioners.  IC0 and IC2 (Prec::ic0 / ic2,
preconditioner_form.cpp:16-80, preconditioner_solve.cpp:14-73) against the
reference's own build (<case>_prec.npz): x after 1 and 5 iterations and at
ResRat 1e-8 bit for bit, the printed iteration counts and residuals.  The
red-black symmetric Gauss-Seidel (the library's parallel replacement) has no
counterpart in the reference; it is pinned by its closed form: z = M^-1 r
with M = (D - L) D^-1 (D - U) assembled densely in red-black order, M's
symmetry, and the classical splitting identity M = A + L D^-1 U.  CPU only."""
import json
import os

import numpy as np
import pytest

import psb_oracle as O
from conftest import GOLDEN, PREC_CASES, load_golden


def load_prec(name):
    z = np.load(os.path.join(GOLDEN, name + "_prec.npz"), allow_pickle=False)
    return {k: z[k] for k in z.files if k != "meta"}, json.loads(str(z["meta"]))


@pytest.mark.parametrize("case", PREC_CASES)
@pytest.mark.parametrize("prec", ["ic0", "ic2"])
def test_incomplete_cholesky_cg_matches_reference(case, prec):
    g = load_golden(case)
    arr, meta = load_prec(case)
    A = O.Mat(*g.shape, g.mat())
    logs = meta["runs"][prec]
    for q, (k, snap) in enumerate(((1, "1"), (5, "5"), (2000, "full"))):
        x, info = O.cg_prec(A, prec, np.zeros(A.n), g["b"], g["dV"], maxit=k, res_rat=1e-8,
                            res_tol=1e-15)
        assert np.array_equal(x, arr[f"{prec}_x{snap}"]), (case, prec, snap)
        res, ratio, iters = logs[q]
        assert info["iters"] == iters and info["res"] == res


def dense(A, shape):
    """The 7-diagonal matrix as a dense array (row P: c_P on the diagonal,
    -a_nb at the neighbours)."""
    nx, ny, nz = shape
    n = nx * ny * nz
    M = np.diag(A["c"].astype(np.float64))
    for P in range(n):
        i, j, k = P // (ny * nz), (P // nz) % ny, P % nz
        for key, di, dj, dk in (("w", -1, 0, 0), ("e", 1, 0, 0), ("s", 0, -1, 0),
                                ("n", 0, 1, 0), ("b", 0, 0, -1), ("t", 0, 0, 1)):
            ii, jj, kk = i + di, j + dj, k + dk
            if 0 <= ii < nx and 0 <= jj < ny and 0 <= kk < nz:
                M[P, (ii * ny + jj) * nz + kk] = -A[key][P]
    return M


def random_system(shape, seed, sym=True):
    nx, ny, nz = shape
    n = nx * ny * nz
    rng = np.random.default_rng(seed)
    A = {k: rng.uniform(0.1, 1.0, n) for k in "wesnbt"}
    v = lambda a: a.reshape(shape)
    v(A["w"])[0] = 0; v(A["e"])[-1] = 0; v(A["s"])[:, 0] = 0; v(A["n"])[:, -1] = 0
    v(A["b"])[:, :, 0] = 0; v(A["t"])[:, :, -1] = 0
    if sym:  # e_P = w_E, n_P = s_N, t_P = b_T
        v(A["e"])[:-1] = v(A["w"])[1:]
        v(A["n"])[:, :-1] = v(A["s"])[:, 1:]
        v(A["t"])[:, :, :-1] = v(A["b"])[:, :, 1:]
    A["c"] = sum(A[k] for k in "wesnbt") + rng.uniform(0.0, 0.3, n)
    return A


@pytest.mark.parametrize("shape,sym", [((4, 3, 5), True), ((5, 4, 3), False), ((3, 3, 3), True)])
def test_rbsgs_is_the_red_black_gauss_seidel_closed_form(shape, sym):
    A = random_system(shape, 11, sym)
    n = int(np.prod(shape))
    Ad = dense(A, shape)
    ii, jj, kk = np.meshgrid(*[np.arange(m) for m in shape], indexing="ij")
    even = ((ii + jj + kk) % 2 == 0).ravel()
    order = np.concatenate([np.where(even)[0], np.where(~even)[0]])
    Ar = Ad[np.ix_(order, order)]  # even cells first
    D = np.diag(np.diag(Ar))
    L, U = -np.tril(Ar, -1), -np.triu(Ar, 1)  # A = D - L - U
    Mr = (D - L) @ np.linalg.inv(D) @ (D - U)
    r = np.random.default_rng(5).standard_normal(n)
    zr = np.linalg.solve(Mr, r[order])
    z = np.empty(n)
    z[order] = zr
    zo = O.prec_apply(O.Mat(*shape, A), "rbsgs", r)
    np.testing.assert_allclose(zo, z, rtol=1e-12, atol=1e-12 * np.abs(z).max())
    # the other colour first is a different preconditioner: the test tells them apart
    order2 = np.concatenate([np.where(~even)[0], np.where(even)[0]])
    A2 = Ad[np.ix_(order2, order2)]
    D2 = np.diag(np.diag(A2))
    M2 = (D2 + np.tril(A2, -1)) @ np.linalg.inv(D2) @ (D2 + np.triu(A2, 1))
    z2 = np.empty(n)
    z2[order2] = np.linalg.solve(M2, r[order2])
    assert np.abs(z2 - z).max() > 1e-6 * np.abs(z).max()
    # splitting identity, and symmetry for a symmetric A
    np.testing.assert_allclose(Mr, Ar + L @ np.linalg.inv(D) @ U, rtol=1e-12, atol=1e-12)
    if sym:
        r2 = np.random.default_rng(6).standard_normal(n)
        z2o = O.prec_apply(O.Mat(*shape, A), "rbsgs", r2)
        assert r2 @ zo == pytest.approx(r @ z2o, rel=1e-12)


def test_rbsgs_without_couplings_is_the_diagonal():
    shape = (4, 4, 4)
    A = random_system(shape, 2)
    for k in "wesnbt":
        A[k][:] = 0.0
    r = np.random.default_rng(1).standard_normal(64)
    M = O.Mat(*shape, A)
    assert np.array_equal(O.prec_apply(M, "rbsgs", r), r / A["c"])


@pytest.mark.parametrize("case", PREC_CASES)
def test_rbsgs_cg_sits_between_diagonal_and_ic2(case):
    """On the fixture systems (density ratio 1000) CG with the red-black
    Gauss-Seidel needs about half the iterations of the diagonal
    preconditioner and 1.5-1.8 times those of the reference's IC2."""
    g = load_golden(case)
    _, meta = load_prec(case)
    A = O.Mat(*g.shape, g.mat())
    it = {}
    for p in ("di", "rbsgs"):
        x, info = O.cg_prec(A, p, np.zeros(A.n), g["b"], g["dV"], maxit=2000, res_rat=1e-8,
                            res_tol=1e-15)
        assert info["converged"]
        it[p] = info["iters"]
    ic2 = meta["runs"]["ic2"][-1][2]
    assert it["rbsgs"] <= 0.6 * it["di"]
    assert ic2 < it["rbsgs"] <= 2.0 * ic2
