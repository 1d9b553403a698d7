"""Pressure::discretize and Pressure::update_rhs on the device (through the C
ABI) against the oracle's restatement of the reference, bit for bit: the
assembled matrix is checked through the bit-exact SpMV (A x on the device
equals the oracle's A x for random x) and by solving with it; the right-hand
side directly; Pressure::ghost's surface-tension source against the
reference's own fext."""
import math

import numpy as np
import pytest

import psb
import psb_oracle as O
from conftest import GHOST_CASES, GOLDEN, case_metrics, load_ghost, load_golden

pytestmark = pytest.mark.gpu


def stretched_cylinder_inputs(nx, ny, nz, beta=2.5, lz=2.0, d=0.6, nsph=6, seed=5):
    h = 1.0 / max(nx, ny, nz)
    xn = np.array([h * i for i in range(nx + 1)])
    yn = np.array([h * i for i in range(ny + 1)])
    tb = math.tanh(beta)
    zn = np.array([0.5 * lz * (1.0 + math.tanh(beta * (2.0 * k / nz - 1.0)) / tb)
                   for k in range(nz + 1)])
    zn[0], zn[-1] = 0.0, lz
    mx, my, mz = O.metrics(xn), O.metrics(yn), O.metrics(zn)
    sph = O.spheres(seed, nsph, 0.02, 0.08, (xn[-1], yn[-1], zn[-1]))
    rho = O.density(mx[0], my[0], mz[0], sph, 1.0, 1000.0)
    R, cx, cz = 0.5 * d * xn[-1], 0.5 * xn[-1], 0.5 * zn[-1]
    off = np.zeros((nx, ny, nz), dtype=np.uint8)
    for i in range(nx):
        for k in range(nz):
            if math.hypot(mx[0][i] - cx, mz[0][k] - cz) < R:
                off[i, :, k] = 1
    return mx, my, mz, rho, off.ravel()


@pytest.mark.parametrize("solid", [False, True])
def test_assembled_matrix_and_rhs_are_the_references(solid):
    nx, ny, nz = 12, 8, 24
    mx, my, mz, rho, off = stretched_cylinder_inputs(nx, ny, nz)
    off = off if solid else None
    A, _, _ = O.discretize(mx[1:], my[1:], mz[1:], rho, off)
    s = psb.Solver(nx, ny, nz)
    s.assemble_pressure(rho, mx[1:], my[1:], mz[1:], off)
    rng = np.random.default_rng(1)
    for _ in range(3):
        x = rng.standard_normal(A.n)
        assert np.array_equal(s.apply(x), O.spmv(A, x))
    # right-hand side from device face velocities
    u, v, w = O.face_velocity(1005, nx, ny, nz)
    du, dv, dw = (psb.DeviceArray.from_host(a) for a in (u, v, w))
    db = psb.DeviceArray(A.n)
    s.update_rhs_device(du.addr, dv.addr, dw.addr, db.addr, zero_mean=False)
    b_ref = O.rhs(mx[1], my[1], mz[1], u, v, w, off=off, zero_mean=False)
    assert np.array_equal(db.download(), b_ref)
    s.update_rhs_device(du.addr, dv.addr, dw.addr, db.addr, zero_mean=True)
    b0 = O.rhs(mx[1], my[1], mz[1], u, v, w, off=off, zero_mean=True)
    b = db.download()
    assert np.max(np.abs(b - b0)) <= 1e-14 * np.max(np.abs(b0))  # mean summed in tree order
    # the assembled system solves like the uploaded one
    x1, r1 = s.solve(np.zeros(A.n), b0, precond=psb.PSB_PREC_AC, norm=psb.PSB_NORM_L2REL,
                     res_rat=1e-9, res_tol=0.0, max_iter=500, ac_coarsen=1, ac_cfac=1.0)
    s2 = psb.Solver(nx, ny, nz)
    if solid:
        s2.set_active(off)
    s2.set_matrix({k: A[k] for k in "cwesnbt"})
    s2.set_geometry(mx[1], my[1], mz[1])
    x2, r2 = s2.solve(np.zeros(A.n), b0, precond=psb.PSB_PREC_AC, norm=psb.PSB_NORM_L2REL,
                      res_rat=1e-9, res_tol=0.0, max_iter=500, ac_coarsen=1, ac_cfac=1.0)
    assert r1["converged"] and r1["iterations"] == r2["iterations"]
    assert np.array_equal(x1, x2)


@pytest.mark.parametrize("solid", [False, True])
@pytest.mark.parametrize("mask", [7, 5])
def test_projection_is_the_references(solid, mask):
    """Momentum::project on the device equals the oracle bit for bit
    (stretched grid, variable density, with and without solid cells)."""
    nx, ny, nz = 12, 8, 24
    mx, my, mz, rho, off = stretched_cylinder_inputs(nx, ny, nz)
    off = off if solid else None
    s = psb.Solver(nx, ny, nz)
    s.assemble_pressure(rho, mx[1:], my[1:], mz[1:], off)
    u, v, w = O.face_velocity(1007, nx, ny, nz)
    p = np.random.default_rng(9).standard_normal(nx * ny * nz)
    du, dv, dw, dp, dr = (psb.DeviceArray.from_host(a) for a in (u, v, w, p, rho))
    s.project_device(dr.addr, dp.addr, du.addr, dv.addr, dw.addr, dt=0.37, mask=mask)
    u1, v1, w1 = O.project(mx[1:], my[1:], mz[1:], rho, p, u, v, w, dt=0.37, off=off, mask=mask)
    assert np.array_equal(du.download(), u1)
    assert np.array_equal(dv.download(), v1)
    assert np.array_equal(dw.download(), w1)


def test_solve_then_project_leaves_the_residual_as_divergence_at_128():
    """Full chain on the device at 128^3 (assembly, right-hand side, mixed
    solve, projection): the divergence of the projected field, formed by
    update_rhs again, is the solve's residual b - A x (a size-independent
    identity), so it is below the solve's tolerance."""
    n = 128
    xn = np.linspace(0.0, 1.0, n + 1)
    m = O.metrics(xn)
    sph = O.spheres(21, 8, 0.05, 0.15)
    rho = O.density(m[0], m[0], m[0], sph, 1.0, 1000.0)
    s = psb.Solver(n, n, n)
    s.assemble_pressure(rho, m[1:], m[1:], m[1:], None)
    u, v, w = O.face_velocity(1009, n, n, n)
    du, dv, dw, dr = (psb.DeviceArray.from_host(a) for a in (u, v, w, rho))
    db, dx = psb.DeviceArray(n ** 3), psb.DeviceArray(n ** 3)
    s.update_rhs_device(du.addr, dv.addr, dw.addr, db.addr, zero_mean=False)
    dx.zero()
    r = s.solve_device(dx.addr, db.addr, precond=psb.PSB_PREC_AC, norm=psb.PSB_NORM_L2REL,
                       res_rat=1e-10, res_tol=0.0, max_iter=500, precision=psb.PSB_MIXED,
                       inner_adapt=2)
    assert r["converged"]
    s.project_device(dr.addr, dx.addr, du.addr, dv.addr, dw.addr)
    ddiv = psb.DeviceArray(n ** 3)
    s.update_rhs_device(du.addr, dv.addr, dw.addr, ddiv.addr, zero_mean=False)
    b, x, div = db.download(), dx.download(), ddiv.download()
    res = s.apply(x, b)
    nb = np.linalg.norm(b)
    assert np.linalg.norm(div - res) <= 1e-12 * nb
    assert np.linalg.norm(div) <= 1.0e-10 * nb * 1.01


def test_checkpoint_warm_start_on_the_device(tmp_path):
    """A converged pressure saved as .bck from the device and loaded into a
    fresh device array starts the next solve converged (0 iterations)."""
    n = 32
    d = psb.generate(n, n, n, 4, 0.05, 0.15, 1.0, 1000.0, 3, 1003)
    s = psb.Solver(n, n, n)
    s.set_matrix(d)
    s.set_geometry(d["dxc"], d["dyc"], d["dzc"])
    db, dx = psb.DeviceArray.from_host(d["rhs"]), psb.DeviceArray(n ** 3)
    dx.zero()
    stop = dict(precond=psb.PSB_PREC_AC, norm=psb.PSB_NORM_L2REL, res_rat=0.0, res_tol=1e-9,
                max_iter=200)
    r = s.solve_device(dx.addr, db.addr, **stop)
    assert r["converged"]
    path = str(tmp_path / psb.bck_name("press", 7, 0))
    s.save_bck_device(path, dx.addr)
    dy = psb.DeviceArray(n ** 3)
    dy.zero()
    s.load_bck_device(path, dy.addr)
    assert np.array_equal(dy.download(), dx.download())
    r2 = s.solve_device(dy.addr, db.addr, **stop)
    assert r2["converged"] and r2["iterations"] <= 1


@pytest.mark.parametrize("case", GHOST_CASES)
def test_ghost_source_is_the_references(case):
    """Pressure::ghost on the device (psb_pressure_ghost_device) against the
    reference's own fext bit for bit (uniform and stretched grid), its
    accumulation into a given fext against the oracle, and solid cells
    skipped."""
    g, z = load_golden(case), load_ghost(case)
    mx, my, mz = case_metrics(g)
    s = psb.Solver(*g.shape)
    s.assemble_pressure(g["rho"], mx, my, mz)
    dc, dk, dr = (psb.DeviceArray.from_host(a) for a in (z["c_g"], z["kappa_g"], g["rho"]))
    n = int(np.prod(g.shape))
    df = psb.DeviceArray.from_host(np.zeros(n))
    sigma = float(z["sigma"])
    s.pressure_ghost_device(dc.addr, dk.addr, dr.addr, sigma, df.addr)
    f1 = df.download()
    assert np.array_equal(f1, z["fext"])
    s.pressure_ghost_device(dc.addr, dk.addr, dr.addr, sigma, df.addr)
    assert np.array_equal(df.download(),
                          O.ghost(mx, my, mz, z["c_g"], z["kappa_g"], g["rho"], sigma, fext=f1))
    off = np.zeros(n, dtype=np.uint8)
    off[np.flatnonzero(f1)[::3]] = 1
    s2 = psb.Solver(*g.shape)
    s2.assemble_pressure(g["rho"], mx, my, mz, off)
    df2 = psb.DeviceArray.from_host(np.zeros(n))
    s2.pressure_ghost_device(dc.addr, dk.addr, dr.addr, sigma, df2.addr)
    assert np.array_equal(df2.download(),
                          O.ghost(mx, my, mz, z["c_g"], z["kappa_g"], g["rho"], sigma, off=off))


def test_boundary_conditions_are_the_references():
    """psb_set_boundary + psb_assemble_pressure with Dirichlet pressure sides
    and a Neumann side with a value: the SpMV of the assembled matrix and the
    right-hand side with the boundary source bit for bit against the
    reference's own build (bnd fixture), and the multigrid-preconditioned CG
    on the now non-singular system (its coarsest level the plain inverse)
    reaches the direct solution."""
    g = load_golden("bubbles16x12x8")
    z = np.load(f"{GOLDEN}/bubbles16x12x8_bnd.npz", allow_pickle=False)
    nx, ny, nz = g.shape
    n = nx * ny * nz
    mx, my, mz = case_metrics(g)
    s = psb.Solver(nx, ny, nz)
    s.set_boundary(z["types"], z["values"])
    s.assemble_pressure(g["rho"], mx, my, mz)
    A = O.Mat(nx, ny, nz, {k: z["A_" + k] for k in "cwesnbt"})
    rng = np.random.default_rng(3)
    for _ in range(2):
        x = rng.standard_normal(n)
        assert np.array_equal(s.apply(x), O.spmv(A, x))
    u, v, w = O.face_velocity(1005, nx, ny, nz)
    du, dv, dw = (psb.DeviceArray.from_host(a) for a in (u, v, w))
    db = psb.DeviceArray(n)
    s.update_rhs_device(du.addr, dv.addr, dw.addr, db.addr, zero_mean=False)
    fb = O.discretize(mx, my, mz, g["rho"], bnd=(z["types"], z["values"]))[3]
    assert np.array_equal(db.download(), O.rhs(mx[0], my[0], mz[0], u, v, w, zero_mean=False,
                                               fbnd=fb))
    s.update_rhs_device(du.addr, dv.addr, dw.addr, db.addr, zero_mean=True)
    b = db.download()
    assert np.max(np.abs(b - z["b"])) <= 1e-14 * np.max(np.abs(z["b"]))
    x, r = s.solve(np.zeros(n), b, precond=psb.PSB_PREC_AC, norm=psb.PSB_NORM_L2REL,
                   res_rat=1e-11, res_tol=0.0, max_iter=200)
    assert r["converged"]
    import scipy.sparse.linalg as spla
    xd = spla.spsolve(dense_csr(A, (nx, ny, nz)), b)
    assert np.linalg.norm(x - xd) <= 1e-8 * np.linalg.norm(xd)
    # the mixed-precision forms and the preconditioned BiCGS reach it too
    for kw in (dict(precision=psb.PSB_MIXED, inner_adapt=2, inner_rtol=1e-2),
               dict(precision=psb.PSB_MIXED, inner_adapt=0, inner_rtol=1e-4),
               dict(precision=psb.PSB_MIXED, inner_adapt=2, inner_rtol=1e-2, ac_kcycle=-1),
               dict(method=psb.PSB_BICGS)):
        xm, rm = s.solve(np.zeros(n), b, precond=psb.PSB_PREC_AC, norm=psb.PSB_NORM_L2REL,
                         res_rat=1e-10, res_tol=0.0, max_iter=300, **kw)
        assert rm["converged"], kw
        assert np.linalg.norm(xm - xd) <= 1e-7 * np.linalg.norm(xd), kw
    # walls again: no boundary source
    s.set_boundary([0] * 6, [0.0] * 6)
    s.assemble_pressure(g["rho"], mx, my, mz)
    s.update_rhs_device(du.addr, dv.addr, dw.addr, db.addr, zero_mean=False)
    assert np.array_equal(db.download(), O.rhs(mx[0], my[0], mz[0], u, v, w, zero_mean=False))


def dense_csr(A, dims):
    import scipy.sparse as sp
    nx, ny, nz = dims
    N = nx * ny * nz
    idx = np.arange(N).reshape(dims)
    rows, cols, vals = [np.arange(N)], [np.arange(N)], [A.a["c"]]
    for key, sh in (("w", (-1, 0, 0)), ("e", (1, 0, 0)), ("s", (0, -1, 0)), ("n", (0, 1, 0)),
                    ("b", (0, 0, -1)), ("t", (0, 0, 1))):
        a = A.a[key].reshape(dims)
        dst = np.roll(idx, -np.array(sh), axis=(0, 1, 2))
        m = a != 0
        rows.append(idx[m]); cols.append(dst[m]); vals.append(-a[m])
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(N, N))
