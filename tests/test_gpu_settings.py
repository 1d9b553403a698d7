"""Solves with different settings in a row on ONE context give what each
gives on a fresh context: the captured graphs are kept per setting (the FP16
smoother copies, fused post-sweeps, the FP16 CG operator, K-cycle levels,
the flexible beta, omega and the coarse-correction factor are all baked into
a graph), so no switch reuses a graph captured for other settings and none
destroys a graph still in use.  Regression test for settings switches on one
context (ac_store / ac_fuse / ac_kcycle changes, and FP64 / mixed solves with
different ac_cfac / ac_omega sharing one context)."""
import numpy as np
import pytest

import psb

pytestmark = pytest.mark.gpu

N = 64
BASE = dict(precond=psb.PSB_PREC_AC, norm=psb.PSB_NORM_L2REL, res_rat=1e-8, res_tol=0.0,
            max_iter=400)
MIXED = dict(precision=psb.PSB_MIXED, inner_adapt=2, inner_rtol=1e-2)
SEQUENCE = [
    ("fp64 cfac 2", dict(precision=psb.PSB_FP64, ac_cfac=2.0)),
    ("mixed cfac 1", dict(MIXED, ac_cfac=1.0)),
    ("fp64 cfac 1", dict(precision=psb.PSB_FP64, ac_cfac=1.0)),
    ("fp64 kcycle", dict(precision=psb.PSB_FP64, ac_kcycle=-1)),
    ("mixed fp16 smoother", dict(MIXED, ac_store=1)),
    ("mixed fp16 smoother kcycle", dict(MIXED, ac_store=1, ac_kcycle=-1)),
    ("mixed fp16 smoother and cg", dict(MIXED, ac_store=2)),
    ("mixed fused post-sweeps", dict(MIXED, ac_store=1, ac_fuse=1)),
    ("mixed fp32 smoother omega 0.8", dict(MIXED, ac_omega=0.8)),
    ("mixed kcycle fletcher-reeves", dict(MIXED, ac_kcycle=-1, ac_flexible=0)),
    ("fp64 cfac 2 again", dict(precision=psb.PSB_FP64, ac_cfac=2.0)),
    ("mixed fp16 smoother kcycle again", dict(MIXED, ac_store=1, ac_kcycle=-1)),
    ("restarted refinement", dict(precision=psb.PSB_MIXED, inner_adapt=0, inner_rtol=1e-4)),
    ("mixed cfac 1 again", dict(MIXED, ac_cfac=1.0)),
]


@pytest.fixture(scope="module")
def system():
    d = psb.generate(N, N, N, 10, 0.05, 0.15, 1.0, 1000.0, 21, 1021)
    return d


def fresh(d, kw):
    s = psb.Solver(N, N, N)
    s.set_matrix(d)
    db = psb.DeviceArray.from_host(d["rhs"])
    dx = psb.DeviceArray(N ** 3)
    dx.zero()
    r = s.solve_device(dx.addr, db.addr, **dict(BASE, **kw))
    x = dx.download()
    s.close()
    return x, r


def test_settings_switches_on_one_context(system):
    d = system
    s = psb.Solver(N, N, N)
    s.set_matrix(d)
    db = psb.DeviceArray.from_host(d["rhs"])
    dx = psb.DeviceArray(N ** 3)
    for name, kw in SEQUENCE:
        print(f"[settings] {name}", flush=True)
        dx.zero()
        r = s.solve_device(dx.addr, db.addr, **dict(BASE, **kw))
        x = dx.download()
        xf, rf = fresh(d, kw)
        assert r["converged"], (name, r)
        assert r["iterations"] == rf["iterations"] and r["outer"] == rf["outer"], (name, r, rf)
        assert np.array_equal(x, xf), (name, np.abs(x - xf).max())


def test_struct_size_is_checked(system):
    d = system
    s = psb.Solver(N, N, N)
    s.set_matrix(d)
    st = psb.stop_defaults(**BASE)
    st.struct_size -= 8  # a caller built against a shorter psb_stop
    r = psb.Result()
    x = np.zeros(N ** 3)
    rc = psb.lib().psb_solve(s._c, psb._dp(x), psb._dp(d["rhs"]), psb.C.byref(st),
                             psb.C.byref(r))
    assert rc == -1
    c = psb._cycle()
    c.struct_size += 4
    assert psb.lib().psb_ac_vcycle(s._c, psb._dp(x), psb._dp(d["rhs"]), psb.C.byref(c),
                                   psb.C.byref(r)) == -1


def test_back_to_back_solves_on_caller_zeroed_arrays(system):
    """Solves in a row on one context, the caller zeroing x with hipMemset
    (null stream) between them: every solve starts from x = 0 (res0 = ||b||
    in the L2 norm) and gives the same result.  A call returns only when its
    last write of x is done, and the context's stream orders after the null
    stream, so neither the memset nor the next solve races the previous
    solve's final update."""
    d = system
    s = psb.Solver(N, N, N)
    s.set_matrix(d)
    db = psb.DeviceArray.from_host(d["rhs"])
    dx = psb.DeviceArray(N ** 3)
    bnorm = np.linalg.norm(d["rhs"])
    for name, kw in (("mixed kcycle", dict(MIXED, ac_store=1, ac_kcycle=-1)),
                     ("fp64 kcycle", dict(precision=psb.PSB_FP64, ac_kcycle=-1))):
        runs = []
        for _ in range(4):
            dx.zero()
            r = s.solve_device(dx.addr, db.addr, **dict(BASE, **kw))
            runs.append((r, dx.download()))
        for r, x in runs:
            assert abs(r["res0"] - bnorm) <= 1e-12 * bnorm, (name, r["res0"], bnorm)
            assert r["iterations"] == runs[0][0]["iterations"], name
            assert np.array_equal(x, runs[0][1]), name
