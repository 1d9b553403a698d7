"""The oracle (oracle/psb_oracle.c) against the reference's own outputs and
against what the mathematics fixes.  CPU only."""
import math

import numpy as np
import pytest
import scipy.sparse as sp

import psb_oracle as O


def dense(A):
    """Dense matrix of the 7-diagonal operator, (Ax)_P = c x_P - sum a_nb x_nb."""
    nx, ny, nz = A.shape
    n = nx * ny * nz
    M = np.zeros((n, n))
    idx = lambda i, j, k: (i * ny + j) * nz + k
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                P = idx(i, j, k)
                M[P, P] = A["c"][P]
                for (di, dj, dk, key) in ((-1, 0, 0, "w"), (1, 0, 0, "e"), (0, -1, 0, "s"),
                                          (0, 1, 0, "n"), (0, 0, -1, "b"), (0, 0, 1, "t")):
                    ii, jj, kk = i + di, j + dj, k + dk
                    if 0 <= ii < nx and 0 <= jj < ny and 0 <= kk < nz:
                        M[P, idx(ii, jj, kk)] = -A[key][P]
    return M


# --------------------------------------------------------------------------
# against the reference (golden fixtures written by oracle/_ref/ref_driver)
# --------------------------------------------------------------------------
def test_discretize_matches_reference(golden):
    mx, my, mz = golden.metrics()
    A, ci, dV = O.discretize(mx, my, mz, golden["rho"])
    for k in "cwesnbt":
        assert np.array_equal(A[k], golden["A_" + k]), k
    assert np.array_equal(ci, golden["A_ci"])
    assert np.array_equal(dV, golden["dV"])


def test_metrics_and_density_match_reference(golden):
    xc, dxc, dxw, dxe = O.metrics(golden["xn"])
    assert np.array_equal(dxc, golden["dxc"]) and np.array_equal(dxw, golden["dxw"])
    assert np.array_equal(dxe, golden["dxe"])
    yc = O.metrics(golden["yn"])[0]
    zc = O.metrics(golden["zn"])[0]
    m = golden.meta
    rho = O.density(xc, yc, zc, golden["spheres"].reshape(-1, 4), m["rho_in"], m["rho_out"])
    assert np.array_equal(rho, golden["rho"])


def test_rhs_matches_reference(golden):
    nx, ny, nz = golden.shape
    u, v, w = O.face_velocity(golden.meta["seed_vel"], nx, ny, nz)
    b = O.rhs(golden["dxc"], golden["dyc"], golden["dzc"], u, v, w)
    assert np.array_equal(b, golden["b"])


def test_spmv_and_residual_bit_exact(golden):
    A = O.Mat(*golden.shape, golden.mat())
    assert np.array_equal(O.spmv(A, golden["x_rand"]), golden["Ax_rand"])
    assert np.array_equal(O.resid(A, golden["b"], golden["x_rand"]), golden["r_rand"])


def test_cg_iterates_match_reference(golden):
    A = O.Mat(*golden.shape, golden.mat())
    b, dV = golden["b"], golden["dV"]
    for k in golden.meta["cg_snapshots"]:
        x, info = O.cg(A, np.zeros(A.n), b, dV, maxit=k, res_rat=1e-8)
        assert np.array_equal(x, golden["cg_x%d" % k])
    x, info = O.cg(A, np.zeros(A.n), b, dV, maxit=2000, res_rat=1e-8)
    assert np.array_equal(x, golden["cg_xfull"])
    res, ratio, iters = golden.info["cg"][-1]
    assert info["iters"] == iters and info["converged"]
    assert info["res"] == pytest.approx(res, rel=1e-15)


def test_bicgs_iterates_match_reference(golden):
    A = O.Mat(*golden.shape, golden.mat())
    b, dV = golden["b"], golden["dV"]
    x, info = O.bicgs(A, np.zeros(A.n), b, dV, maxit=1, res_rat=1e-8)
    assert np.array_equal(x, golden["bicgs_x1"])
    x, info = O.bicgs(A, np.zeros(A.n), b, dV, maxit=2000, res_rat=1e-8)
    assert np.array_equal(x, golden["bicgs_xfull"])
    res, ratio, iters = golden.info["bicgs"][-1]
    assert info["iters"] == iters
    assert info["res"] == pytest.approx(res, rel=1e-15)


def test_jacobi_matches_reference(golden):
    A = O.Mat(*golden.shape, golden.mat())
    x, info = O.jacobi(A, np.zeros(A.n), golden["b"], maxit=5, res_rat=1e-30, res_tol=0.0)
    assert np.array_equal(x, golden["jacobi_x5"])


def test_levels_and_coarse_systems_match_reference(golden):
    A = O.Mat(*golden.shape, golden.mat())
    ac = O.AC(A, golden["dV"], rule=0)
    assert [list(ac.dims(l)) for l in range(ac.nlev)] == golden.info["levels"]
    for l in range(1, ac.nlev):
        m = ac.mat(l)
        for k in "cwesnbt":
            assert np.array_equal(m[k], golden["L%d_%s" % (l, k)]), (l, k)
    for l in range(ac.nlev):
        assert np.array_equal(ac.aflag(l), golden["L%d_aflag" % l])


def test_restriction_and_interpolation_match_reference(golden):
    A = O.Mat(*golden.shape, golden.mat())
    ac = O.AC(A, golden["dV"], rule=0)
    ac.phi(0)[:] = golden["restrict_phi"]
    ac.f(0)[:] = golden["b"]
    ac.restriction(0)
    assert np.array_equal(ac.f(1), golden["restrict_fH"])
    ac.phi(0)[:] = 0.0
    ac.phi(1)[:] = golden["interp_phiH"]
    ac.interpolation(1)
    assert np.array_equal(ac.phi(0), golden["interp_phih"])


def test_ac_vcycle_matches_reference(golden):
    A = O.Mat(*golden.shape, golden.mat())
    ac = O.AC(A, golden["dV"], rule=0)
    phi, info = ac.vcycle(np.zeros(A.n), golden["b"], factor=1e-30, max_cyc=1)
    assert np.array_equal(phi, golden["ac_phi1"])
    phi, info = ac.vcycle(np.zeros(A.n), golden["b"], factor=1e-8, max_cyc=20)
    assert np.array_equal(phi, golden["ac_phifull"])
    ref_full = [r for r in golden.info["acret"] if r[0] == "full"][0]
    assert info["ncyc"] == ref_full[2] and info["converged"] == bool(ref_full[1])
    # per-cycle residuals of the converged run (the last block of "Cycle" lines)
    cyc = golden.info["cycles"][-ref_full[2]:]
    for c, res in cyc:
        assert info["hist"][c] == pytest.approx(res, rel=1e-15)


# --------------------------------------------------------------------------
# closed forms, brute force and invariants
# --------------------------------------------------------------------------
def random_mat(nx, ny, nz, seed, sym=False):
    rng = np.random.default_rng(seed)
    n = nx * ny * nz
    A = {k: rng.uniform(0.1, 1.0, n) for k in "wesnbt"}
    g = lambda a: a.reshape(nx, ny, nz)
    g(A["w"])[0] = 0; g(A["e"])[-1] = 0
    g(A["s"])[:, 0] = 0; g(A["n"])[:, -1] = 0
    g(A["b"])[:, :, 0] = 0; g(A["t"])[:, :, -1] = 0
    if sym:
        g(A["e"])[:-1] = g(A["w"])[1:]
        g(A["n"])[:, :-1] = g(A["s"])[:, 1:]
        g(A["t"])[:, :, :-1] = g(A["b"])[:, :, 1:]
    A["c"] = sum(A[k] for k in "wesnbt") + rng.uniform(0.0, 0.1, n)
    return A


def test_spmv_against_dense_brute_force():
    nx, ny, nz = 5, 4, 3
    A = random_mat(nx, ny, nz, 1)
    M = dense(O.Mat(nx, ny, nz, A))
    x = np.random.default_rng(2).standard_normal(nx * ny * nz)
    Am = O.Mat(nx, ny, nz, A)
    np.testing.assert_allclose(O.spmv(Am, x), M @ x, rtol=1e-14, atol=1e-14)
    b = np.random.default_rng(3).standard_normal(nx * ny * nz)
    np.testing.assert_allclose(O.resid(Am, b, x), b - M @ x, rtol=1e-14, atol=1e-14)


def test_pressure_matrix_invariants():
    A, ci, dV, b = O.sphere_case(10, 3, 0.1, 0.3, 1.0, 1000.0, 4, 1004)
    nx = ny = nz = 10
    g = lambda a: a.reshape(nx, ny, nz)
    # symmetry bit for bit
    assert np.array_equal(g(A["e"])[:-1], g(A["w"])[1:])
    assert np.array_equal(g(A["n"])[:, :-1], g(A["s"])[:, 1:])
    assert np.array_equal(g(A["t"])[:, :, :-1], g(A["b"])[:, :, 1:])
    # walls zero, constants in the null space, sum b = 0
    assert np.all(g(A["w"])[0] == 0) and np.all(g(A["t"])[:, :, -1] == 0)
    r = O.spmv(A, np.ones(A.n))
    assert np.max(np.abs(r)) < 1e-12 * np.max(A["c"])
    assert abs(b.sum()) < 1e-12 * np.abs(b).sum()
    # x^T A x > 0 for x orthogonal to constants
    x = np.random.default_rng(5).standard_normal(A.n)
    x -= x.mean()
    assert x @ O.spmv(A, x) > 0
    # uniform grid closed form away from spheres: a = h/rho
    m = O.Mat(*A.shape, A.a)
    assert np.min(A["w"][A["w"] > 0]) == pytest.approx(0.1 / 1000.0, rel=1e-12)
    assert np.max(A["w"]) == pytest.approx(0.1 / 1.0, rel=1e-12)


def test_dots_against_fsum():
    rng = np.random.default_rng(7)
    x, y = rng.standard_normal(1000), rng.standard_normal(1000)
    dV = rng.uniform(0.5, 2.0, 1000)
    assert O.dot(x, y) == pytest.approx(math.fsum(x * y), rel=1e-13)
    ref = math.fsum(x * y / dV / dV) / 1000.0 / 1000.0
    assert O.dot_voldiv_avg(x, y, dV) == pytest.approx(ref, rel=1e-13)


def pinv_solution(A, b):
    M = dense(A)
    x = np.linalg.lstsq(M, b, rcond=None)[0]
    return x - x.mean()


def test_cg_and_bicgs_reach_dense_solution():
    A, ci, dV, b = O.sphere_case(4, 1, 0.3, 0.3, 1.0, 100.0, 1, 11)
    xs = pinv_solution(A, b)
    x, info = O.cg(A, np.zeros(A.n), b, dV, maxit=500, res_rat=1e-14, res_tol=0.0)
    assert info["converged"]
    np.testing.assert_allclose(x - x.mean(), xs, atol=1e-10 * np.abs(xs).max())
    x, info = O.bicgs(A, np.zeros(A.n), b, dV, maxit=500, res_rat=1e-13, res_tol=0.0)
    np.testing.assert_allclose(x - x.mean(), xs, atol=1e-8 * np.abs(xs).max())


def test_cg_search_directions_are_a_orthogonal_in_exact_steps():
    # two CG steps from x0 = 0: x1 in span{z0}, x2 in span{z0, z1}; the residual
    # after each step is orthogonal to the preconditioned residuals before it
    A, ci, dV, b = O.sphere_case(6, 2, 0.15, 0.3, 1.0, 10.0, 3, 13)
    x1, _ = O.cg(A, np.zeros(A.n), b, dV, maxit=1, res_rat=1e-30, res_tol=0.0)
    r1 = O.resid(A, b, x1)
    z0 = b / A["c"]
    assert abs(r1 @ z0) < 1e-12 * np.linalg.norm(r1) * np.linalg.norm(z0)


def test_jacobi_on_diagonal_matrix_is_exact():
    n = 4 * 3 * 2
    A = {k: np.zeros(n) for k in "wesnbt"}
    A["c"] = np.linspace(1.0, 3.0, n)
    b = np.linspace(-1.0, 2.0, n)
    Am = O.Mat(4, 3, 2, A)
    x, info = O.jacobi(Am, np.zeros(n), b, maxit=1, res_rat=1e-30, res_tol=0.0)
    np.testing.assert_array_equal(x, (1.0 / A["c"]) * b)


def test_smallest_factor_and_level_rules():
    for n in range(2, 200):
        spf = next(p for p in range(2, n + 1) if n % p == 0)
        assert O.lib().ora_smallest_factor(n) == spf
    assert O.levels(0, 64, 64, 64) == [(64,) * 3, (32,) * 3, (16,) * 3, (8,) * 3, (4,) * 3]
    assert O.levels(0, 384, 384, 768)[-1] == (3, 3, 3)
    assert O.levels(1, 256, 256, 256)[-1] == (8, 8, 8) and len(O.levels(1, 256, 256, 256)) == 6
    assert O.levels(1, 48, 128, 64) == [(48, 128, 64), (24, 64, 32), (12, 32, 16),
                                        (12, 16, 8), (12, 8, 8)]
    assert O.levels(1, 384, 384, 768)[-1] == (12, 12, 12)


def prolongation(fine, coarse):
    nx, ny, nz = fine
    X, Y, Z = coarse
    CI, CJ, CK = nx // X, ny // Y, nz // Z
    rows, cols = [], []
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                rows.append((i * ny + j) * nz + k)
                cols.append(((i // CI) * Y + j // CJ) * Z + k // CK)
    return sp.csr_matrix((np.ones(len(rows)), (rows, cols)), shape=(nx * ny * nz, X * Y * Z))


def test_galerkin_restriction_prolongation_closed_form():
    A, ci, dV, b = O.sphere_case(8, 2, 0.15, 0.3, 1.0, 1000.0, 6, 16)
    ac = O.AC(A, dV, rule=0)
    assert ac.dims(1) == (4, 4, 4)
    P = prolongation((8, 8, 8), (4, 4, 4)).toarray()
    M = dense(A)
    MH = dense(O.Mat(4, 4, 4, ac.mat(1)))
    np.testing.assert_allclose(MH, P.T @ M @ P, rtol=1e-12, atol=1e-14)
    phi = np.random.default_rng(9).standard_normal(A.n)
    ac.phi(0)[:] = phi
    ac.f(0)[:] = b
    ac.restriction(0)
    rH = P.T @ (b - M @ phi)
    rH -= rH.mean()
    np.testing.assert_allclose(ac.f(1), rH, rtol=1e-11, atol=1e-13 * np.abs(rH).max())
    eH = np.random.default_rng(10).standard_normal(64)
    ac.phi(0)[:] = phi
    ac.phi(1)[:] = eH
    ac.interpolation(1)
    np.testing.assert_allclose(ac.phi(0), phi + P @ eH, rtol=1e-15)


def test_ac_preconditioner_is_symmetric_and_exact_on_coarsest():
    A, ci, dV, b = O.sphere_case(16, 4, 0.1, 0.25, 1.0, 1000.0, 8, 18)
    ac = O.AC(A, dV, rule=1)
    assert ac.nlev == 2
    rng = np.random.default_rng(11)
    u, v = rng.standard_normal(A.n), rng.standard_normal(A.n)
    u -= u.mean(); v -= v.mean()
    Mu, Mv = ac.precond(u), ac.precond(v)
    assert v @ Mu == pytest.approx(u @ Mv, rel=1e-10)
    assert u @ Mu > 0
    # coarsest operator is the pseudo-inverse: A_H G r = r for r of zero mean
    G = ac.coarse_dense()
    MH = dense(O.Mat(*ac.dims(1), ac.mat(1)))
    r = rng.standard_normal(G.shape[0]); r -= r.mean()
    np.testing.assert_allclose(MH @ (G @ r), r, atol=1e-9 * np.abs(r).max())
    assert abs((G @ r).sum()) < 1e-10 * np.abs(G @ r).sum()


def test_acpcg_reaches_dense_solution():
    A, ci, dV, b = O.sphere_case(16, 4, 0.1, 0.25, 1.0, 1000.0, 8, 18)
    ac = O.AC(A, dV, rule=1)
    x, info = ac.acpcg(np.zeros(A.n), b, rtol=1e-12, maxit=200)
    assert info["converged"] and info["iters"] < 40
    xs = pinv_solution(A, b)
    np.testing.assert_allclose(x - x.mean(), xs, atol=1e-9 * np.abs(xs).max())


def test_planted_mistake_is_caught():
    """A sign slip in one coupling (planted in a copy of the oracle's matrix)
    breaks the reference fixture comparison and the dense check."""
    from conftest import load_golden
    g = load_golden("drop12")
    A = dict(g.mat())
    A["s"] = A["s"].copy()
    A["s"][100] = -A["s"][100]
    Am = O.Mat(*g.shape, A)
    assert not np.array_equal(O.spmv(Am, g["x_rand"]), g["Ax_rand"])


def test_zline_solve_against_dense_tridiagonal_solves():
    """The z-line smoother's column solve equals a dense solve of each
    column's tridiagonal block (diagonal c, off-diagonals -b)."""
    rng = np.random.default_rng(7)
    nx, ny, nz = 3, 4, 9
    n = nx * ny * nz
    arr = {k: np.zeros(n) for k in "cwesnbt"}
    b = rng.uniform(0.1, 1.0, n).reshape(nx, ny, nz)
    b[:, :, 0] = 0.0
    arr["b"] = b.ravel()
    tb = np.zeros((nx, ny, nz))
    tb[:, :, :-1] = b[:, :, 1:]
    arr["t"] = tb.ravel()
    arr["c"] = (b + tb + rng.uniform(0.5, 2.0, (nx, ny, nz))).ravel()
    A = O.Mat(nx, ny, nz, arr)
    t = rng.standard_normal(n)
    x = O.zline_solve(A, t).reshape(nx, ny, nz)
    c3, t3 = arr["c"].reshape(nx, ny, nz), t.reshape(nx, ny, nz)
    for i in range(nx):
        for j in range(ny):
            T = np.diag(c3[i, j]) - np.diag(b[i, j, 1:], -1) - np.diag(b[i, j, 1:], 1)
            assert np.allclose(x[i, j], np.linalg.solve(T, t3[i, j]), rtol=1e-12, atol=1e-14)


def test_zline_preconditioner_is_symmetric():
    A, _, _, _ = O.sphere_case(16, 3, 0.1, 0.25, 1.0, 100.0, 4, 1004)
    ac = O.AC(A, np.ones(A.n), rule=1)
    rng = np.random.default_rng(3)
    u, v = rng.standard_normal(A.n), rng.standard_normal(A.n)
    u -= u.mean()
    v -= v.mean()
    mu = ac.precond(u, omega=0.9, cfac=2.0, smoother=1)
    mv = ac.precond(v, omega=0.9, cfac=2.0, smoother=1)
    assert abs(u @ mv - v @ mu) <= 1e-10 * abs(u @ mv)
