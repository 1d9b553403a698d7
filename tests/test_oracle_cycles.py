"""AC::cycle in the oracle against the reference's own build: the F and W
cycles, full-cycle starts (fullcycle), the linf control with the reference's
stale previous-linf value, the divergence check with and without
priority_min_cyc, min_cycles, per-stage MaxIter and stale counts
(stale_iter_array) and the level solves' ResRat -- phi after each run bit for bit, every logged residual (17
significant digits) and the returned cycle count; and six of those runs with
BiCGS as every level's solver (the reference's AC with a non-symmetric
equation's solver), on the case's matrix and on a non-symmetric one.  CPU
only."""
import numpy as np
import pytest

import psb_oracle as O
from conftest import (CYCLE_CASES, CYCLE_KINDS, FLEX_CASES, cycle_kwargs, cycle_matrix,
                      load_cycles, load_golden)


def _runs():
    out = []
    for case in CYCLE_CASES:
        for kind in CYCLE_KINDS:
            _, runs = load_cycles(case, kind)
            out += [(case, kind, r["tag"]) for r in runs]
    return out


@pytest.fixture(scope="module")
def systems():
    res = {}
    for case in CYCLE_CASES:
        g = load_golden(case)
        res[case] = {kind: (g, load_cycles(case, kind)) for kind in CYCLE_KINDS}
    return res


@pytest.mark.parametrize("case,kind,tag", _runs())
def test_cycle_matches_reference(systems, case, kind, tag):
    g, (phis, runs) = systems[case][kind]
    run = [r for r in runs if r["tag"] == tag][0]
    A = O.Mat(*g.shape, cycle_matrix(g, kind, phis))
    ac = O.AC(A, g["dV"], rule=0)
    phi, info = ac.cycle(np.zeros(A.n), g["b"], **cycle_kwargs(run))
    assert np.array_equal(phi, phis[tag + "_phi"])
    assert info["ncyc"] == run["ncyc"] and info["converged"] == run["converged"]
    assert info["hist"][0] == pytest.approx(run["initial"][0], rel=1e-15)
    assert info["hist_linf"][0] == pytest.approx(run["initial"][1], rel=1e-15)
    assert len(info["hist"]) == 1 + len(run["cycles"])
    for q, (c, res, linf) in enumerate(run["cycles"]):
        assert info["hist"][q + 1] == pytest.approx(res, rel=1e-15), (c, q)
        assert info["hist_linf"][q + 1] == pytest.approx(linf, rel=1e-15), (c, q)


def test_fixtures_exercise_the_controls():
    """The runs cover what they are meant to: a divergence stop, a run that
    goes on past it without the check, min_cycles deferring it, a full-cycle
    start whose result the divergence check discards, F and W shapes that
    converge, and the linf control."""
    _, runs = load_cycles("bubbles16x12x8")
    r = {x["tag"]: x for x in runs}
    assert r["Vdiv"]["ncyc"] < r["Vdiv"]["max_cyc"] and not r["Vdiv"]["converged"]
    assert r["Vnodiv"]["ncyc"] == r["Vnodiv"]["max_cyc"]
    assert r["Vmin"]["ncyc"] >= r["Vmin"]["min_cyc"]
    assert r["Ffull"]["converged"] and r["Wfull"]["converged"]
    assert r["FMVfull"]["converged"] and r["FMVfull"]["cycles"][0][0] == 0
    assert r["Vlinf"]["converged"] and r["Vlinf"]["use_linf"]


@pytest.mark.parametrize("tag", ["Vstale", "Fstale"])
def test_stale_counts_change_the_result(systems, tag):
    """The stale-count runs differ from the same settings with the counts off,
    so their fixtures pin where each count goes (pre-, re-, post-smoothing)."""
    g, (phis, runs) = systems["bubbles16x12x8"]["cycles"]
    run = [r for r in runs if r["tag"] == tag][0]
    kw = cycle_kwargs(run)
    A = O.Mat(*g.shape, g.mat())
    kw["ms"] = (-1, -1, -1)
    phi, _ = O.AC(A, g["dV"], rule=0).cycle(np.zeros(A.n), g["b"], **kw)
    assert not np.array_equal(phi, phis[tag + "_phi"])
    if tag == "Fstale":  # the re-smoothing count alone swapped with the others
        kw["ms"] = (4, 2, 4)
        phi, _ = O.AC(A, g["dV"], rule=0).cycle(np.zeros(A.n), g["b"], **kw)
        assert not np.array_equal(phi, phis[tag + "_phi"])


def _flex_runs():
    return [(case, r["tag"]) for case in FLEX_CASES for r in load_cycles(case, "flex")[1]]


@pytest.mark.parametrize("case,tag", _flex_runs())
def test_flex_cycle_matches_reference(case, tag):
    """AC::flexcycle (additive_recursive.cpp:161-200) where the reference's
    recursion ends: phi bit for bit, the logged residuals, the cycle count
    and the return value; the bound on level visits is not reached."""
    g = load_golden(case)
    phis, runs = load_cycles(case, "flex")
    run = [r for r in runs if r["tag"] == tag][0]
    A = O.Mat(*g.shape, g.mat())
    ac = O.AC(A, g["dV"], rule=0)
    phi, info = ac.cycle(np.zeros(A.n), g["b"], **cycle_kwargs(run))
    assert not info["bound"]
    assert np.array_equal(phi, phis[tag + "_phi"])
    assert info["ncyc"] == run["ncyc"] and info["converged"] == run["converged"]
    assert len(info["hist"]) == 1 + len(run["cycles"])
    for q, (c, res, linf) in enumerate(run["cycles"]):
        assert info["hist"][q + 1] == pytest.approx(res, rel=1e-15), (c, q)


def test_flex_cycle_bound_ends_a_run_the_reference_does_not_end():
    """On bubbles16 with the level solves' ResRat 0.1 the reference's flex
    recursion does not return (its smoother on one level cannot meet the
    ratio, so it goes down and up between two levels); the oracle's bound on
    level visits ends the run as a divergence would: phi of the last
    completed cycle, not converged."""
    g = load_golden("bubbles16")
    A = O.Mat(*g.shape, g.mat())
    ac = O.AC(A, g["dV"], rule=0)
    phi, info = ac.cycle(np.zeros(A.n), g["b"], loop="flex", factor=1e-8, max_cyc=20,
                         mv=(20, 20, 20), resrat=0.1, max_visits=400)
    assert info["bound"] and not info["converged"]
    assert np.all(np.isfinite(phi))


def test_bicgs_level_solves_change_the_result(systems):
    """The BiCGS runs are not the CG runs under another name: the same
    settings with CG level solves give another phi (so the fixtures pin the
    level solver), and the oracle's BiCGS itself matches the reference's
    single-level fixtures elsewhere (test_oracle)."""
    g, (phis, runs) = systems["bubbles16x12x8"]["cycles_bicgs"]
    A = O.Mat(*g.shape, g.mat())
    for run in runs:
        kw = dict(cycle_kwargs(run), solver="cg")
        phi, _ = O.AC(A, g["dV"], rule=0).cycle(np.zeros(A.n), g["b"], **kw)
        assert not np.array_equal(phi, phis[run["tag"] + "_phi"]), run["tag"]


@pytest.mark.parametrize("case", CYCLE_CASES)
def test_nonsymmetric_fixture_and_its_levels(case):
    """The non-symmetric runs' matrix is non-symmetric on every level the
    oracle coarsens (east coupling of a cell against the west coupling of
    its neighbour), and its row sums exceed the couplings (the time term)."""
    g = load_golden(case)
    phis, _ = load_cycles(case, "cycles_nonsym")
    M = cycle_matrix(g, "cycles_nonsym", phis)
    A = O.Mat(*g.shape, M)
    ac = O.AC(A, g["dV"], rule=0)
    for l in range(ac.nlev - 1):
        nx, ny, nz = ac.dims(l)
        m = {k: np.asarray(v).reshape(nx, ny, nz) for k, v in ac.mat(l).items()}
        assert np.abs(m["e"][:-1] - m["w"][1:]).max() > 1e-3 * np.abs(m["w"]).max(), l
        off = sum(m[k] for k in "wesnbt")
        assert np.all(m["c"] > off)
