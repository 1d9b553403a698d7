"""Pressure::ghost (pressure_ghost.cpp:10-190), the ghost-fluid method's
surface-tension source, in the oracle against the reference's own build
(ghost fixtures: the colour function of the case, a seeded curvature field
with zeros and both signs, on a uniform and on a stretched grid) bit for bit,
and against what the method fixes: on a uniform grid every crossed face adds
equal and opposite amounts to its two cells, so fext sums to zero face by
face.  CPU only."""
import numpy as np
import pytest

import psb_oracle as O
from conftest import GHOST_CASES, case_metrics as metrics, load_ghost, load_golden


@pytest.mark.parametrize("case", GHOST_CASES)
def test_ghost_source_matches_reference(case):
    g, z = load_golden(case), load_ghost(case)
    assert np.array_equal(z["rho"], g["rho"]) and np.array_equal(z["dxc"], g["dxc"])
    f = O.ghost(*metrics(g), z["c_g"], z["kappa_g"], g["rho"], float(z["sigma"]))
    assert np.array_equal(f, z["fext"])
    assert np.count_nonzero(f) > 50
    # fext accumulates: a second call adds the same terms again
    f2 = O.ghost(*metrics(g), z["c_g"], z["kappa_g"], g["rho"], float(z["sigma"]), fext=f)
    assert np.allclose(f2, 2.0 * f, rtol=1e-14, atol=0)


def test_ghost_source_cancels_across_faces_on_a_uniform_grid():
    g, z = load_golden("bubbles16x12x8"), load_ghost("bubbles16x12x8")
    nx, ny, nz = g.shape
    f = O.ghost(*metrics(g), z["c_g"], z["kappa_g"], g["rho"], float(z["sigma"]))
    assert abs(f.sum()) <= 1e-14 * np.abs(f).sum()
    # only cells next to the c = 0.5 crossing receive a source
    c = z["c_g"].reshape(nx + 2, ny + 2, nz + 2)
    ci = c[1:-1, 1:-1, 1:-1]
    near = np.zeros_like(ci, dtype=bool)
    for ax in range(3):
        for sh in (-1, 1):
            nb = np.roll(c, sh, axis=ax)[1:-1, 1:-1, 1:-1]
            near |= (ci - 0.5) * (nb - 0.5) < 0
    assert np.all(f.reshape(nx, ny, nz)[~near] == 0.0)
    # solid cells are skipped
    off = np.zeros(f.size, dtype=np.uint8)
    off[np.flatnonzero(f)[::3]] = 1
    fo = O.ghost(*metrics(g), z["c_g"], z["kappa_g"], g["rho"], float(z["sigma"]), off=off)
    assert np.all(fo[off == 1] == 0.0) and np.array_equal(fo[off == 0], f[off == 0])
