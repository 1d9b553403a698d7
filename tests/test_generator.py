"""The product's case generator (psb_generate, host C++) against the oracle,
bit for bit, including the stretched grid and the solid cylinder.  CPU only:
psb_generate makes no HIP call."""
import math

import numpy as np
import pytest

import psb
import psb_oracle as O


@pytest.mark.parametrize("shape,nsph,rmin,rmax,rin,rout,centre", [
    ((16, 16, 16), 1, 0.25, 0.25, 1000.0, 1.0, (0.5, 0.5, 0.5)),
    ((24, 16, 20), 5, 0.05, 0.2, 1.0, 1000.0, None),
    ((32, 32, 32), 16, 0.03, 0.12, 1.0, 1000.0, None),
    ((12, 20, 9), 3, 0.05, 0.3, 10.0, 1.0, None),
])
def test_generator_matches_oracle(shape, nsph, rmin, rmax, rin, rout, centre):
    nx, ny, nz = shape
    g = psb.generate(nx, ny, nz, nsph, rmin, rmax, rin, rout, 2, 1002, centre=centre)
    A, ci, dV, b = O.sphere_case(nx, nsph, rmin, rmax, rin, rout, 2, 1002, centre=centre,
                                 ny=ny, nz=nz)
    for k in "cwesnbt":
        assert np.array_equal(g[k], A[k]), k
    assert np.array_equal(g["rhs"], b)
    assert np.array_equal(np.repeat(g["dxc"], ny * nz) * np.tile(np.repeat(g["dyc"], nz), nx) *
                          np.tile(g["dzc"], nx * ny), dV)


def test_generator_stretched_with_cylinder_matches_oracle():
    nx, ny, nz, beta, lz, d = 12, 8, 24, 2.5, 2.0, 0.6
    g = psb.generate(nx, ny, nz, 6, 0.02, 0.08, 1.0, 1000.0, 5, 1005, stretch_z=beta, lz=lz,
                     cylinder=True, cyl_d=d)
    h = 1.0 / max(nx, ny, nz)
    xn = np.array([h * i for i in range(nx + 1)])
    yn = np.array([h * i for i in range(ny + 1)])
    tb = math.tanh(beta)
    zn = np.array([0.5 * lz * (1.0 + math.tanh(beta * (2.0 * k / nz - 1.0)) / tb)
                   for k in range(nz + 1)])
    zn[0], zn[-1] = 0.0, lz
    mx, my, mz = O.metrics(xn), O.metrics(yn), O.metrics(zn)
    sph = O.spheres(5, 6, 0.02, 0.08, (xn[-1], yn[-1], zn[-1]))
    rho = O.density(mx[0], my[0], mz[0], sph, 1.0, 1000.0)
    R, cx, cz = 0.5 * d * xn[-1], 0.5 * xn[-1], 0.5 * zn[-1]
    off = np.zeros((nx, ny, nz), dtype=np.uint8)
    for i in range(nx):
        for k in range(nz):
            dx, dz = mx[0][i] - cx, mz[0][k] - cz
            if math.sqrt(dx * dx + dz * dz) < R:
                off[i, :, k] = 1
    assert off.sum() > 0
    A, ci, dV = O.discretize(mx[1:], my[1:], mz[1:], rho, off)
    for k in "cwesnbt":
        assert np.array_equal(g[k], A[k]), k
    assert np.array_equal(g["off"], off.ravel())
    u, v, w = O.face_velocity(1005, nx, ny, nz)
    b = O.rhs(mx[1], my[1], mz[1], u, v, w, off=off)
    assert np.array_equal(g["rhs"], b)
    # decoupled solid rows, zero RHS there
    o = off.ravel().astype(bool)
    for k in "wesnbt":
        assert np.all(g[k][o] == 0.0)
    assert np.all(g["rhs"][o] == 0.0)
    assert abs(g["rhs"][~o].sum()) < 1e-12 * np.abs(g["rhs"]).sum()
