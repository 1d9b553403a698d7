"""The oracle's BiCGS with the V-cycle as its preconditioner (not a reference
configuration: the reference's BiCGS takes a diagonal or IC preconditioner),
pinned by what the method fixes.  Right-preconditioned BiCGStab has the
residuals of plain BiCGStab (H. A. van der Vorst, SIAM J. Sci. Stat. Comput.
13 (1992) 631-644) on the operator A M^-1, with x = x0 + M^-1 y, and the
V-cycle M^-1 (a fixed number of Jacobi sweeps from zero, block-sum
restriction, the dense coarsest inverse, constant prolongation) is linear:
a textbook BiCGStab in numpy on the dense A M^-1 of a small system gives the
oracle's residual history and iterate.  The diagonal form of the same loop
is pinned on the reference's BiCGS vectors (test_oracle.py).  CPU only."""
import numpy as np
import pytest

import psb_oracle as O

KW = dict(nu1=2, nu2=2, omega=0.9, cfac=2.0)


def case(n=16, seed=6, ny=8, nz=8):
    A, ci, dV, b = O.sphere_case(n, 3, 0.1, 0.25, 1.0, 1000.0, seed, 100 + seed, ny=ny, nz=nz)
    return A, np.ones(b.size), b


def dense_minv(ac, n):
    """Columns M^-1 e_j of the V-cycle on level 0."""
    M = np.empty((n, n))
    e = np.zeros(n)
    for j in range(n):
        e[j] = 1.0
        M[:, j] = ac.precond(e, **KW)
        e[j] = 0.0
    return M


def bicgstab_right(A, Minv, b, x0, iters):
    """Plain BiCGStab on B = A M^-1 from y = 0; returns the residual norms
    and x0 + M^-1 y after each iteration."""
    B = A @ Minv
    r = b - A @ x0
    rt = r.copy()
    y = np.zeros_like(b)
    p = np.zeros_like(b)
    v = np.zeros_like(b)
    rho_old = alpha = omega = 1.0
    hist, xs = [np.linalg.norm(r)], []
    for k in range(iters):
        rho = rt @ r
        p = r.copy() if k == 0 else r + (rho / rho_old) * (alpha / omega) * (p - omega * v)
        v = B @ p
        alpha = rho / (rt @ v)
        s = r - alpha * v
        t = B @ s
        omega = (t @ s) / (t @ t)
        y = y + alpha * p + omega * s
        r = s - omega * t
        rho_old = rho
        hist.append(np.linalg.norm(r))
        xs.append(x0 + Minv @ y)
    return np.array(hist), xs


def dense_A(M, n):
    x = np.zeros(n)
    A = np.empty((n, n))
    for j in range(n):
        x[j] = 1.0
        A[:, j] = O.spmv(M, x)
        x[j] = 0.0
    return A


def test_acbicgs_is_bicgstab_on_the_right_preconditioned_operator():
    A, dV, b = case()
    n = b.size
    ac = O.AC(A, dV, rule=1)
    assert ac.nlev >= 2
    Minv = dense_minv(ac, n)
    Ad = dense_A(A, n)
    # the cycle is linear: its dense form reproduces an application
    rng = np.random.default_rng(1)
    w = rng.standard_normal(n)
    assert np.allclose(Minv @ w, ac.precond(w, **KW), rtol=0, atol=1e-12 * np.abs(Minv @ w).max())
    K = 6
    h_ref, xs = bicgstab_right(Ad, Minv, b, np.zeros(n), K)
    for k in (1, 3, K):
        x, info = ac.acbicgs(np.zeros(n), b, dV, maxit=k, res_rat=0.0, res_tol=0.0, **KW)
        assert info["reason"] == 0 and info["iters"] == k + 1
        # the oracle's norm with unit volumes: sqrt(r.r) / N
        h = info["hist"] * n
        assert np.allclose(h, h_ref[:k + 1], rtol=1e-8, atol=0), (k, h, h_ref[:k + 1])
        assert np.linalg.norm(x - xs[k - 1]) <= 1e-8 * np.linalg.norm(xs[k - 1])


def test_acbicgs_converges_to_the_cg_solution():
    A, dV, b = case(16, seed=3, ny=16, nz=16)
    n = b.size
    ac = O.AC(A, dV, rule=1)
    x, info = ac.acbicgs(np.zeros(n), b, dV, maxit=200, res_rat=1e-10, res_tol=0.0, **KW)
    assert info["converged"] and info["reason"] == 2
    assert info["iters"] < 40
    # the recursive residual is the true one to rounding
    r = b - O.spmv(A, x)
    assert np.linalg.norm(r) < 1e-9 * np.linalg.norm(b)
    xc, ic = ac.acpcg(np.zeros(n), b, rtol=1e-12, maxit=200, **KW)
    assert ic["converged"]
    d = (x - x.mean()) - (xc - xc.mean())
    assert np.linalg.norm(d) < 1e-7 * np.linalg.norm(xc - xc.mean())
    # fewer iterations than the diagonal BiCGS on the same system
    _, idg = O.bicgs(A, np.zeros(n), b, dV, maxit=2000, res_rat=1e-10, res_tol=0.0)
    assert info["iters"] * 4 < idg["iters"]


def test_acbicgs_small_s_exit_and_zero_rhs():
    A, dV, b = case()
    n = b.size
    ac = O.AC(A, dV, rule=1)
    # b = 0 from x = 0: the early return of bicgs.cpp:70 (reason 7)
    x, info = ac.acbicgs(np.zeros(n), np.zeros(n), dV, maxit=5, **KW)
    assert info["reason"] == 7 and info["iters"] == 0 and np.all(x == 0)
    # |s| below res_tol after the first half step (bicgs.cpp:134-138): the
    # loop ends with x = alpha M^-1 r (reason 5); a small scale keeps the
    # initial residual above res_tol
    z = ac.precond(b, **KW)
    v = O.spmv(A, z)
    alpha = (b @ b) / (b @ v)
    ns = np.linalg.norm(b - alpha * v)
    tol = np.sqrt(ns * np.linalg.norm(b))
    assert ns < tol < np.linalg.norm(b)
    x, info = ac.acbicgs(np.zeros(n), b, dV, maxit=5, res_rat=0.0, res_tol=tol, scale=1e-9,
                         **KW)
    assert info["reason"] == 5 and info["iters"] == 1 and not info["converged"]
    assert np.linalg.norm(x - alpha * z) <= 1e-12 * np.linalg.norm(x)


def nonsym(A, shape):
    """The pattern the reference driver uses for its non-symmetric cycle
    runs: east and top couplings scaled per cell, the diagonal raised by the
    added coupling and by 5 % of itself."""
    nx, ny, nz = shape
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    fe = (0.025 * ((7 * i + 3 * j + 5 * k) % 11)).ravel()
    ft = (0.25 / 6.0 * ((2 * i + 5 * j + 3 * k) % 7)).ravel()
    M = {q: np.array(A.a[q], dtype=np.float64) for q in "cwesnbt"}
    de, dt = fe * M["e"], ft * M["t"]
    M["e"] = M["e"] + de
    M["t"] = M["t"] + dt
    M["c"] = M["c"] + de + dt
    M["c"] = M["c"] + 0.05 * M["c"]
    return O.Mat(nx, ny, nz, M)


def test_nonsymmetric_coarsest_inverse_and_right_preconditioning():
    """A non-symmetric matrix: the coarsest operator is the inverse of that
    level's matrix (G A = I), and the preconditioned BiCGS is again plain
    BiCGStab on the dense A M^-1."""
    A0, dV, b = case()
    A = nonsym(A0, (16, 8, 8))
    n = b.size
    ac = O.AC(A, dV, rule=1)
    G = ac.coarse_dense()
    nx, ny, nz = ac.dims(ac.nlev - 1)
    Ac = dense_A(O.Mat(nx, ny, nz, ac.mat(ac.nlev - 1)), nx * ny * nz)
    assert np.abs(Ac - Ac.T).max() > 1e-3 * np.abs(Ac).max()
    assert np.abs(G @ Ac - np.eye(nx * ny * nz)).max() < 1e-12
    Minv = dense_minv(ac, n)
    Ad = dense_A(A, n)
    K = 6
    h_ref, xs = bicgstab_right(Ad, Minv, b, np.zeros(n), K)
    x, info = ac.acbicgs(np.zeros(n), b, dV, maxit=K, res_rat=0.0, res_tol=0.0, **KW)
    assert np.allclose(info["hist"] * n, h_ref, rtol=1e-8, atol=0)
    assert np.linalg.norm(x - xs[-1]) <= 1e-8 * np.linalg.norm(xs[-1])
    # converged: the dense solution of the non-singular system
    x, info = ac.acbicgs(np.zeros(n), b, dV, maxit=100, res_rat=1e-11, res_tol=0.0, **KW)
    assert info["converged"] and info["iters"] < 30
    xd = np.linalg.solve(Ad, b)
    assert np.linalg.norm(x - xd) < 1e-9 * np.linalg.norm(xd)


def test_singular_nonsymmetric_coarsest_level_is_reported():
    """Without the time term a non-symmetric Neumann matrix is singular on
    every level: the coarsest inverse reports it instead of dividing by a
    vanishing pivot."""
    A0, dV, b = case()
    M = {q: np.array(A0.a[q], dtype=np.float64) for q in "cwesnbt"}
    nx, ny, nz = 16, 8, 8
    i = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")[0].ravel()
    de = 0.2 * (i % 3) * M["e"]
    M["e"] = M["e"] + de
    M["c"] = M["c"] + de  # row sums stay zero: the constant is a right null vector
    ac = O.AC(O.Mat(nx, ny, nz, M), dV, rule=1)
    with pytest.raises(ValueError):
        ac.coarse_dense()
    with pytest.raises(ValueError):
        ac.acbicgs(np.zeros(b.size), b, dV, maxit=5, **KW)
