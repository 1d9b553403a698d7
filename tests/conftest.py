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
GOLDEN_CASES = ["drop12", "bubbles16x12x8", "bubbles16"]


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
