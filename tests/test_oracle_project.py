"""Momentum::project in the oracle (momentum_project.cpp:10-76).

The projection is pinned by the discrete identity that makes it the
pressure-correction step: for ANY pressure p, the right-hand side that
Pressure::update_rhs forms from the corrected velocities equals the old one
minus dt * (A p), with A the matrix Pressure::discretize builds from the same
density and metrics, on every fluid cell.  Both sides come from independently
pinned oracle functions (ora_rhs, ora_discretize, ora_spmv), and the grids
are stretched in all three directions with a variable density, so a wrong
sign, neighbour, face density or centre distance (dxc for dxw) breaks it.
Solving A p = b(u) and projecting therefore leaves a velocity field whose
divergence is the solver's residual."""
import math
import sys
import os

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import psb_oracle as O  # noqa: E402
from conftest import GOLDEN, case_metrics, load_golden  # noqa: E402


def stretched(n, beta, seed, length=1.0):
    rng = np.random.default_rng(seed)
    tb = math.tanh(beta)
    xn = np.array([0.5 * length * (1.0 + math.tanh(beta * (2.0 * q / n - 1.0)) / tb)
                   for q in range(n + 1)])
    xn[1:-1] += rng.uniform(-0.1, 0.1, n - 1) * (length / n) * 0.5
    xn[0], xn[-1] = 0.0, length
    return xn


def case(nx=10, ny=9, nz=12, solid=False, seed=3):
    xn, yn, zn = stretched(nx, 1.5, seed), stretched(ny, 2.0, seed + 1), \
        stretched(nz, 2.5, seed + 2, 2.0)
    mx, my, mz = O.metrics(xn), O.metrics(yn), O.metrics(zn)
    sph = O.spheres(seed, 4, 0.1, 0.3, (xn[-1], yn[-1], zn[-1]))
    rho = O.density(mx[0], my[0], mz[0], sph, 1.0, 1000.0)
    off = None
    if solid:
        off = np.zeros((nx, ny, nz), dtype=np.uint8)
        off[3:6, 2:5, 4:9] = 1
        off = off.ravel()
    return mx, my, mz, rho, off


@pytest.mark.parametrize("solid", [False, True])
@pytest.mark.parametrize("dt", [1.0, 0.37])
def test_projection_removes_dt_times_A_p_from_the_divergence(solid, dt):
    mx, my, mz, rho, off = case(solid=solid)
    nx, ny, nz = len(mx[1]), len(my[1]), len(mz[1])
    A, _, _ = O.discretize(mx[1:], my[1:], mz[1:], rho, off)
    u, v, w = O.face_velocity(11, nx, ny, nz)
    p = np.random.default_rng(2).standard_normal(A.n)
    b0 = O.rhs(mx[1], my[1], mz[1], u, v, w, off=off, zero_mean=False)
    u1, v1, w1 = O.project(mx[1:], my[1:], mz[1:], rho, p, u, v, w, dt=dt, off=off)
    b1 = O.rhs(mx[1], my[1], mz[1], u1, v1, w1, off=off, zero_mean=False)
    want = b0 - dt * O.spmv(A, p)
    fluid = np.ones(A.n, bool) if off is None else off == 0
    scale = np.max(np.abs(b0)) + dt * np.max(np.abs(O.spmv(A, np.abs(p))))
    assert np.max(np.abs(b1 - want)[fluid]) <= 1e-13 * scale
    if off is not None:
        assert np.all(b1[~fluid] == 0.0)


def test_wall_faces_and_unselected_components_stay():
    mx, my, mz, rho, _ = case()
    nx, ny, nz = len(mx[1]), len(my[1]), len(mz[1])
    u, v, w = O.face_velocity(12, nx, ny, nz)
    p = np.random.default_rng(4).standard_normal(nx * ny * nz)
    u1, v1, w1 = O.project(mx[1:], my[1:], mz[1:], rho, p, u, v, w, mask=1)
    U0, U1 = u.reshape(nx + 1, ny, nz), u1.reshape(nx + 1, ny, nz)
    assert np.array_equal(U0[0], U1[0]) and np.array_equal(U0[nx], U1[nx])
    assert np.all(U0[1:nx] != U1[1:nx])
    assert np.array_equal(v, v1) and np.array_equal(w, w1)
    _, v2, w2 = O.project(mx[1:], my[1:], mz[1:], rho, p, u, v, w, mask=6)
    V0, V2 = v.reshape(nx, ny + 1, nz), v2.reshape(nx, ny + 1, nz)
    W0, W2 = w.reshape(nx, ny, nz + 1), w2.reshape(nx, ny, nz + 1)
    assert np.array_equal(V0[:, [0, ny]], V2[:, [0, ny]])
    assert np.array_equal(W0[:, :, [0, nz]], W2[:, :, [0, nz]])


def test_solid_faces_untouched_and_both_forms_agree_without_solids():
    mx, my, mz, rho, off = case(solid=True)
    nx, ny, nz = len(mx[1]), len(my[1]), len(mz[1])
    u, v, w = O.face_velocity(13, nx, ny, nz)
    p = np.random.default_rng(5).standard_normal(nx * ny * nz)
    u1, _, _ = O.project(mx[1:], my[1:], mz[1:], rho, p, u, v, w, off=off, mask=1)
    o = off.reshape(nx, ny, nz).astype(bool)
    touch = np.zeros((nx + 1, ny, nz), bool)
    touch[1:nx] = o[1:] | o[:-1]
    assert np.array_equal(u.reshape(nx + 1, ny, nz)[touch], u1.reshape(nx + 1, ny, nz)[touch])
    # an all-fluid body mask takes the body form: same values up to rounding
    ua, va, wa = O.project(mx[1:], my[1:], mz[1:], rho, p, u, v, w)
    ub, vb, wb = O.project(mx[1:], my[1:], mz[1:], rho, p, u, v, w,
                           off=np.zeros(nx * ny * nz, np.uint8))
    for a, b in ((ua, ub), (va, vb), (wa, wb)):
        assert np.allclose(a, b, rtol=1e-14, atol=1e-14)


def test_projected_field_is_divergence_free_to_the_solver_residual():
    mx, my, mz, rho, _ = case(8, 8, 8)
    nx, ny, nz = 8, 8, 8
    A, _, dV = O.discretize(mx[1:], my[1:], mz[1:], rho)
    u, v, w = O.face_velocity(14, nx, ny, nz)
    b = O.rhs(mx[1], my[1], mz[1], u, v, w, zero_mean=False)
    assert abs(b.sum()) <= 1e-12 * np.abs(b).sum()  # closed walls: compatible
    x, info = O.cg(A, np.zeros(A.n), b, dV, mini=1, maxit=2000, res_rat=1e-13, res_tol=0.0)
    assert info["converged"]
    u1, v1, w1 = O.project(mx[1:], my[1:], mz[1:], rho, x, u, v, w)
    div = O.rhs(mx[1], my[1], mz[1], u1, v1, w1, zero_mean=False)
    r = O.resid(A, b, x)
    assert np.max(np.abs(div - r)) <= 1e-12 * np.max(np.abs(b))
    assert np.max(np.abs(div)) <= 1e-9 * np.max(np.abs(b))


def test_boundary_conditions_match_reference():
    """Centered::create_system_bnd with Dirichlet pressure sides (imin 0.3,
    kmax -0.2) and a Neumann side with a value (jmax 0.05): the oracle's
    matrix, boundary source fbnd and right-hand side bit for bit against the
    reference's own build; the Dirichlet sides keep their coupling in c, so
    the matrix is non-singular (no zero row sums there)."""
    g = load_golden("bubbles16x12x8")
    z = np.load(os.path.join(GOLDEN, "bubbles16x12x8_bnd.npz"), allow_pickle=False)
    mx, my, mz = case_metrics(g)
    A, _, _, fb = O.discretize(mx, my, mz, g["rho"], bnd=(z["types"], z["values"]))
    for k in "cwesnbt":
        assert np.array_equal(A.a[k], z["A_" + k]), k
    assert np.array_equal(fb, z["fbnd"])
    nx, ny, nz = g.shape
    u, v, w = O.face_velocity(1005, nx, ny, nz)
    b = O.rhs(mx[0], my[0], mz[0], u, v, w, zero_mean=True, fbnd=fb)
    assert np.array_equal(b, z["b"])
    rows = A.a["c"] - sum(A.a[k] for k in "wesnbt")
    c = rows.reshape(nx, ny, nz)
    assert np.all(c[0] > 0) and np.all(c[:, :, -1] > 0)
    assert np.all(np.abs(c[1:-1, :, :-1]) <= 1e-12 * A.a["c"].max())
