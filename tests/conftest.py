import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("oracle", "psi-boil_amd"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")
GOLDEN_CASES = ["drop12", "bubbles16x12x8", "bubbles16", "stretch12x10x8"]
PREC_CASES = ["drop12", "bubbles16x12x8", "bubbles16"]  # <case>_prec.npz: IC0 / IC2 runs
CYCLE_CASES = ["bubbles16x12x8", "bubbles16"]  # <case>_cycles.npz: AC::cycle runs


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libpsb.so)")


class Golden:
    def __init__(self, name):
        z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
        self.a = {k: z[k] for k in z.files if k != "meta"}
        self.meta = json.loads(str(z["meta"]))
        self.shape = tuple(self.meta["shape"])
        self.info = self.meta["info"]

    def __getitem__(self, k):
        return self.a[k]

    def mat(self, prefix="A_"):
        return {k: self.a[prefix + k] for k in "cwesnbt"}

    def metrics(self):
        return ((self["dxc"], self["dxw"], self["dxe"]), (self["dyc"], self["dys"], self["dyn"]),
                (self["dzc"], self["dzb"], self["dzt"]))


@pytest.fixture(scope="session", params=GOLDEN_CASES)
def golden(request):
    return Golden(request.param)


def load_golden(name):
    return Golden(name)


FLEX_CASES = ["bubbles16x12x8", "bubbles16"]  # <case>_flex.npz: AC::flexcycle runs


def load_cycles(name, kind="cycles"):
    """(phi of each run by tag, the runs' settings and logs) of <name>_<kind>.npz
    (kind "cycles", "cycles_bicgs", "cycles_nonsym" or "flex")."""
    z = np.load(os.path.join(GOLDEN, name + "_" + kind + ".npz"), allow_pickle=False)
    meta = json.loads(str(z["meta"]))
    return {k: z[k] for k in z.files if k != "meta"}, meta["runs"]


GHOST_CASES = ["bubbles16x12x8", "stretch12x10x8"]  # <case>_ghost.npz: Pressure::ghost runs


def load_ghost(name):
    """Colour, curvature (ghosted), sigma and the reference's fext of one
    Pressure::ghost run."""
    z = np.load(os.path.join(GOLDEN, name + "_ghost.npz"), allow_pickle=False)
    return {k: z[k] for k in z.files}


def case_metrics(g):
    """(dxc, dxw, dxe), (dyc, dys, dyn), (dzc, dzb, dzt) of a golden case."""
    return (g["dxc"], g["dxw"], g["dxe"]), (g["dyc"], g["dys"], g["dyn"]), \
        (g["dzc"], g["dzb"], g["dzt"])


CYCLE_KINDS = ("cycles", "cycles_bicgs", "cycles_nonsym")


def cycle_matrix(g, kind, arrays):
    """The seven coefficient arrays the runs of a cycles fixture used: the
    case's own, or for "cycles_nonsym" the non-symmetric matrix the fixture
    holds."""
    if kind == "cycles_nonsym":
        return {k: arrays["N_A_" + k] for k in "cwesnbt"}
    return g.mat()


def cycle_kwargs(run):
    """The keyword arguments of the oracle's AC.cycle for one reference run."""
    shapes = ["V", "F", "W", "flex"]
    return dict(loop=shapes[run["kind"] % 10], init="V" if run["kind"] >= 10 else None,
                toler=-1.0, factor=run["factor"], max_cyc=run["max_cyc"],
                min_cyc=run["min_cyc"], mv=tuple(run["mv"]), stop_if_div=run["stop_if_div"],
                priority_min_cyc=run["priority_min_cyc"], use_linf=run["use_linf"],
                resrat=run["resrat"], ms=tuple(run.get("ms", (-1, -1, -1))),
                solver="bicgs" if run["tag"].startswith("B") else "cg")
