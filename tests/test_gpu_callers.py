"""The other callers of Linear::solve on the same kernels (momentum_solve.cpp:
69-87 for the three staggered velocity components, Enthalpy, Concentration):
their matrices are 7-diagonal on the staggered shapes (nx+1) x ny x nz,
nx x (ny+1) x nz, nx x ny x (nz+1), with a time term on the diagonal, and
non-symmetric once implicit convection or non-uniform coefficients enter.
CG and BiCGS with the diagonal preconditioner on such systems against the
oracle (the reference's algorithm, pinned on its own fixtures): the first
iterates to rounding, the converged counts and solutions."""
import numpy as np
import pytest

import psb
import psb_oracle as O

pytestmark = pytest.mark.gpu


def staggered_system(nx, ny, nz, seed, sym=True, dt_term=2.0):
    """Diffusion couplings with a time term on the diagonal; sym=False scales
    the east / north / top couplings (an upwind-like convection part)."""
    rng = np.random.default_rng(seed)
    n = nx * ny * nz
    g = lambda a: a.reshape(nx, ny, nz)
    A = {k: np.zeros(n) for k in "cwesnbt"}
    for key, ax in (("w", 0), ("s", 1), ("b", 2)):
        a = rng.uniform(0.2, 1.2, n)
        sl = [slice(None)] * 3
        sl[ax] = 0
        g(a)[tuple(sl)] = 0.0
        A[key] = a
    # e_P = w_{P+I} etc. (the same face), zero on the upper walls
    g(A["e"])[:-1] = g(A["w"])[1:]
    g(A["n"])[:, :-1] = g(A["s"])[:, 1:]
    g(A["t"])[:, :, :-1] = g(A["b"])[:, :, 1:]
    if not sym:
        for key in "ent":
            A[key] *= rng.uniform(1.0, 1.6, n)
    A["c"] = A["w"] + A["e"] + A["s"] + A["n"] + A["b"] + A["t"] + dt_term * rng.uniform(0.5, 1.5, n)
    return A


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("shape", [(17, 12, 10), (16, 13, 10), (16, 12, 11)])
def test_cg_on_staggered_shapes_matches_oracle(shape):
    nx, ny, nz = shape
    A = staggered_system(nx, ny, nz, 5)
    n = nx * ny * nz
    b = np.random.default_rng(6).standard_normal(n)
    dV = np.full(n, 1.0)
    s = psb.Solver(nx, ny, nz)
    s.set_matrix(A)
    Ao = O.Mat(nx, ny, nz, A)
    x3o, _ = O.cg(Ao, np.zeros(n), b, dV, maxit=3, res_rat=1e-12, res_tol=0.0)
    x3, _ = s.solve(np.zeros(n), b, max_iter=3, res_rat=1e-12, res_tol=0.0)
    assert rel(x3, x3o) < 1e-13
    xo, io = O.cg(Ao, np.zeros(n), b, dV, maxit=500, res_rat=1e-10, res_tol=0.0)
    x, r = s.solve(np.zeros(n), b, max_iter=500, res_rat=1e-10, res_tol=0.0)
    assert r["converged"] and io["converged"]
    assert abs(r["iterations"] - io["iters"]) <= 1
    assert rel(x, xo) < 1e-9


@pytest.mark.parametrize("shape", [(17, 12, 10), (16, 12, 11)])
def test_bicgs_on_nonsymmetric_staggered_systems_matches_oracle(shape):
    nx, ny, nz = shape
    A = staggered_system(nx, ny, nz, 7, sym=False)
    n = nx * ny * nz
    b = np.random.default_rng(8).standard_normal(n)
    dV = np.full(n, 1.0)
    s = psb.Solver(nx, ny, nz)
    s.set_matrix(A)
    Ao = O.Mat(nx, ny, nz, A)
    assert np.array_equal(s.apply(b), O.spmv(Ao, b))  # the 7-array operator, bit for bit
    x2o, _ = O.bicgs(Ao, np.zeros(n), b, dV, maxit=2, res_rat=1e-12, res_tol=0.0)
    x2, _ = s.solve(np.zeros(n), b, max_iter=2, res_rat=1e-12, res_tol=0.0,
                    method=psb.PSB_BICGS)
    assert rel(x2, x2o) < 1e-12
    xo, io = O.bicgs(Ao, np.zeros(n), b, dV, maxit=500, res_rat=1e-10, res_tol=0.0)
    x, r = s.solve(np.zeros(n), b, max_iter=500, res_rat=1e-10, res_tol=0.0,
                   method=psb.PSB_BICGS)
    assert r["converged"] and io["converged"]
    assert abs(r["iterations"] - io["iters"]) <= max(2, 0.1 * io["iters"])
    assert rel(x, xo) < 1e-8
    assert rel(s.apply(x), b) < 1e-9
