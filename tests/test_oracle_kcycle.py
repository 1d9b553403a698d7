"""The oracle's K-cycle (two flexible-CG steps as the coarse correction of
chosen levels, Notay and Vassilevski 2008) against an independent numpy
restatement built on the oracle's own level matrices (pinned against the
reference's Galerkin operators elsewhere), and the properties the method
fixes: the preconditioner is homogeneous of degree one, the plain V-cycle
comes back with no level chosen, and the residual the two steps leave on
their level is orthogonal to both search directions.  CPU only."""
import numpy as np
import pytest
import scipy.sparse as sp

import psb_oracle as O


def case(n=64, seed=4):
    A, ci, dV, b = O.sphere_case(n, 5, 0.08, 0.2, 1.0, 1000.0, seed, 100 + seed)
    return A, dV, b


def csr(M, dims):
    nx, ny, nz = dims
    N = nx * ny * nz
    idx = np.arange(N).reshape(dims)
    rows, cols, vals = [np.arange(N)], [np.arange(N)], [M["c"]]
    for key, sh in (("w", (-1, 0, 0)), ("e", (1, 0, 0)), ("s", (0, -1, 0)), ("n", (0, 1, 0)),
                    ("b", (0, 0, -1)), ("t", (0, 0, 1))):
        a = M[key].reshape(dims)
        dst = np.roll(idx, -np.array(sh), axis=(0, 1, 2))
        m = a != 0
        rows.append(idx[m]); cols.append(dst[m]); vals.append(-a[m])
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(N, N))


class Restated:
    """The oracle's cycle (damped Jacobi nu1 / nu2, residual block sums over
    active cells, dense coarsest operator, prolongation cfac * eH) and its
    K-cycle step, in numpy."""

    def __init__(self, ac):
        self.L = []
        for l in range(ac.nlev):
            d = ac.dims(l)
            M = ac.mat(l)
            self.L.append(dict(dims=d, A=csr(M, d), c=np.array(M["c"]),
                               fl=np.array(ac.aflag(l)) != 0))
        self.G = ac.coarse_dense()

    def agg(self, l):
        f, cL = self.L[l]["dims"], self.L[l + 1]["dims"]
        fac = [a // b for a, b in zip(f, cL)]
        i, j, k = np.meshgrid(*[np.arange(n) // q for n, q in zip(f, fac)], indexing="ij")
        return ((i * cL[1] + j) * cL[2] + k).ravel()

    def cycle(self, l, r, nu1, nu2, om, cfac, mask):
        L = self.L[l]
        if l == len(self.L) - 1:
            return self.G @ r
        A, c = L["A"], L["c"]
        z = om * r / c
        for _ in range(nu1 - 1):
            z = z + om * (r - A @ z) / c
        t = np.where(L["fl"], r - A @ z, 0.0)
        ag = self.agg(l)
        rH = np.bincount(ag, weights=t, minlength=len(self.L[l + 1]["c"]))
        rH[~self.L[l + 1]["fl"]] = 0.0
        if l + 1 < len(self.L) - 1 and (mask >> (l + 1)) & 1:
            eH, cf = self.kstep(l + 1, rH, nu1, nu2, om, cfac, mask), 1.0
        else:
            eH, cf = self.cycle(l + 1, rH, nu1, nu2, om, cfac, mask), cfac
        z = z + np.where(L["fl"], cf * eH[ag], 0.0)
        for _ in range(nu2):
            z = z + om * (r - A @ z) / c
        return z

    def kstep(self, l, r, nu1, nu2, om, cfac, mask, parts=False):
        A = self.L[l]["A"]
        c = self.cycle(l, r, nu1, nu2, om, cfac, mask)
        v = A @ c
        rho1 = c @ v
        a1 = (c @ r) / rho1
        rt = r - a1 * v
        d = self.cycle(l, rt, nu1, nu2, om, cfac, mask)
        w = A @ d
        gam, bet, dl = d @ v, d @ w, d @ rt
        a2 = dl / (bet - gam * gam / rho1)
        out = (a1 - gam * a2 / rho1) * c + a2 * d
        return (out, c, d) if parts else out


@pytest.fixture(scope="module")
def setup():
    A, dV, b = case()
    ac = O.AC(A, dV, rule=1)
    assert ac.nlev >= 4
    return A, ac, b, Restated(ac)


@pytest.mark.parametrize("mask", [0, 0b10, 0b110])
def test_kcycle_matches_the_numpy_restatement(setup, mask):
    A, ac, b, R = setup
    r = np.random.default_rng(1).standard_normal(A.n)
    r -= r.mean()
    kw = dict(nu1=2, nu2=2, omega=0.9, cfac=2.0)
    try:
        O.set_kcycle(mask)
        z = ac.precond(r, **kw)
    finally:
        O.set_kcycle(0)
    zr = R.cycle(0, r, 2, 2, 0.9, 2.0, mask)
    assert np.linalg.norm(z - zr) <= 1e-10 * np.linalg.norm(zr)
    if mask:  # the K-cycle is a different preconditioner
        z0 = ac.precond(r, **kw)
        assert np.linalg.norm(z - z0) > 1e-3 * np.linalg.norm(z0)


def test_kcycle_is_homogeneous_and_its_steps_are_galerkin(setup):
    A, ac, b, R = setup
    r = np.random.default_rng(2).standard_normal(A.n)
    r -= r.mean()
    try:
        O.set_kcycle(0b110)
        z1 = ac.precond(r, nu1=2, nu2=2, omega=0.9, cfac=2.0)
        z2 = ac.precond(-3.0 * r, nu1=2, nu2=2, omega=0.9, cfac=2.0)
    finally:
        O.set_kcycle(0)
    np.testing.assert_allclose(z2, -3.0 * z1, rtol=1e-12, atol=1e-12 * np.abs(z1).max())
    # the two flexible-CG steps on level 1 leave a residual orthogonal to both
    # directions (the A-norm error is minimal over their span)
    rH = np.random.default_rng(3).standard_normal(len(R.L[1]["c"]))
    rH[~R.L[1]["fl"]] = 0.0
    out, c, d = R.kstep(1, rH, 2, 2, 0.9, 2.0, 0b110, parts=True)
    res = rH - R.L[1]["A"] @ out
    scale = np.linalg.norm(rH)
    assert abs(c @ res) <= 1e-10 * scale * np.linalg.norm(c)
    assert abs(d @ res) <= 1e-10 * scale * np.linalg.norm(d)


def test_kcycle_preconditioned_cg_converges():
    """On a 48^3 bubble system to 1e-8 CG with the K-cycle on levels 1 and 2
    converges in no more iterations than with the V-cycle (the gain grows with
    the grid: at 256^3 the restatement needs 21 against 36)."""
    A, ci, dV, b = O.sphere_case(48, 8, 0.05, 0.15, 1.0, 1000.0, 7, 107)
    ac = O.AC(A, dV, rule=1)
    kw = dict(rtol=1e-8, maxit=300, nu1=2, nu2=2, omega=0.9, cfac=2.0)
    x0, i0 = ac.acpcg(np.zeros(A.n), b, **kw)
    try:
        O.set_kcycle(0b110)
        x1, i1 = ac.acpcg(np.zeros(A.n), b, **kw)
    finally:
        O.set_kcycle(0)
    assert i0["converged"] and i1["converged"]
    assert i1["iters"] <= i0["iters"], (i0["iters"], i1["iters"])
    r = b - O.spmv(A, x1)
    assert np.linalg.norm(r - r.mean()) < 1e-7 * np.linalg.norm(b - b.mean())


def _orth_run(ac, b, mask, flex):
    kw = dict(rtol=1e-10, maxit=300, nu1=2, nu2=2, omega=0.9, cfac=2.0)
    try:
        O.set_kcycle(mask)
        O.set_flexible(flex)
        return ac.acpcg(np.zeros(b.size), b, orth=True, **kw)
    finally:
        O.set_kcycle(0)
        O.set_flexible(False)


def test_flexible_beta_keeps_directions_a_orthogonal():
    """The K-cycle changes the preconditioner from one application to the
    next, so the Fletcher-Reeves beta rho_new / rho no longer makes the new
    search direction A-orthogonal to the last one; the flexible beta
    -z.A p / p.A p does by construction.  The record is |p_k+1.A p_k| over
    the A-norms of both directions, computed with a separate product."""
    A, ci, dV, b = O.sphere_case(48, 8, 0.05, 0.15, 1.0, 1000.0, 7, 107)
    ac = O.AC(A, dV, rule=1)
    xf, inf = _orth_run(ac, b, 0b110, True)
    xr, inr = _orth_run(ac, b, 0b110, False)
    assert inf["converged"] and inr["converged"]
    assert inf["orth"].size > 5 and inr["orth"].size > 5
    assert inf["orth"].max() < 1e-9, inf["orth"].max()
    assert inr["orth"].max() > 1e-5, inr["orth"].max()
    # both reach the same solution
    r = b - O.spmv(A, xf)
    assert np.linalg.norm(r - r.mean()) < 1e-9 * np.linalg.norm(b - b.mean())
    # with a fixed (linear, symmetric) preconditioner, the V-cycle, the two
    # formulas agree and both keep the directions A-orthogonal to rounding
    xv, iv = _orth_run(ac, b, 0, True)
    xw, iw = _orth_run(ac, b, 0, False)
    assert abs(iv["iters"] - iw["iters"]) <= 1
    assert iv["orth"].max() < 1e-9 and iw["orth"].max() < 1e-6
    assert np.linalg.norm((xv - xv.mean()) - (xw - xw.mean())) < 1e-8 * np.linalg.norm(xv)
