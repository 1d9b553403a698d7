"""psb_solve_device_static (CG with the diagonal preconditioner, no
allocation: its work arrays live in the library's static device array)
gives what psb_solve_device gives, bit for bit, on two calls in a row, the
second on a smaller system; and rejects a system above its size limit."""
import numpy as np
import pytest

import psb

pytestmark = pytest.mark.gpu

STOP = dict(precision=psb.PSB_FP64, max_iter=1000, res_rat=1e-8)


def _pair(n, seed):
    d = psb.generate(n, n, n, 3, 0.1, 0.25, 1000.0, 1.0, seed, 1000 + seed)
    s = psb.Solver(n, n, n)
    s.set_matrix(d)
    s.set_geometry(d["dxc"], d["dyc"], d["dzc"])
    db = psb.DeviceArray.from_host(d["rhs"])
    x0 = np.random.default_rng(seed).standard_normal(n ** 3) * 1e-3
    da, dr = psb.DeviceArray.from_host(x0), psb.DeviceArray.from_host(x0)
    print(f"[no-alloc] {n}^3 seed {seed}", flush=True)
    rs = s.solve_device_static(da.addr, db.addr, **STOP)
    rr = s.solve_device(dr.addr, db.addr, **STOP)
    xs, xr = da.download(), dr.download()
    s.close()
    return rs, rr, xs, xr


def test_static_matches_solve_device_twice():
    for n, seed in ((40, 3), (24, 4)):  # the second call on a smaller system
        rs, rr, xs, xr = _pair(n, seed)
        assert rs["converged"] and rs["status"] == psb.PSB_OK, rs
        for k in ("iterations", "converged", "status", "res0", "res"):
            assert rs[k] == rr[k], (n, k, rs[k], rr[k])
        assert np.array_equal(xs, xr), (n, np.abs(xs - xr).max())


def test_static_rejects_other_methods_and_large_systems():
    n = 16
    d = psb.generate(n, n, n, 1, 0.2, 0.2, 1000.0, 1.0, 0, 1000)
    s = psb.Solver(n, n, n)
    s.set_matrix(d)
    db = psb.DeviceArray.from_host(d["rhs"])
    dx = psb.DeviceArray(n ** 3)
    dx.zero()
    with pytest.raises(psb.PsbError):
        s.solve_device_static(dx.addr, db.addr, **dict(STOP, precond=psb.PSB_PREC_AC))
    s.close()
    nx, ny, nz = 160, 128, 128  # 2.6 M cells, above the 2^21-cell limit
    big = psb.generate(nx, ny, nz, 1, 0.2, 0.2, 1000.0, 1.0, 0, 1000)
    s = psb.Solver(nx, ny, nz)
    s.set_matrix(big)
    db = psb.DeviceArray.from_host(big["rhs"])
    dx = psb.DeviceArray(nx * ny * nz)
    dx.zero()
    with pytest.raises(psb.PsbError, match="invalid argument"):
        s.solve_device_static(dx.addr, db.addr, **STOP)
    s.close()
