"""Batch case generator and batched-ABI argument checks (CPU only), and the
oracle's level rule of the batched hierarchy (known answers)."""
import ctypes as C

import psb
import psb_oracle as O


def test_batch_cases_are_seeded_and_in_range():
    a = psb.batch_cases(4, 96)
    assert a == psb.batch_cases(4, 96)
    assert a != psb.batch_cases(5, 96)
    dims = {32, 48, 64, 96, 128}
    for c in a:
        assert {c["nx"], c["ny"], c["nz"]} <= dims
        assert 1 <= c["nsph"] <= 8
        side = min(c["nx"], c["ny"], c["nz"]) / max(c["nx"], c["ny"], c["nz"])
        assert abs(c["rmin"] - 0.05 * side) < 1e-15 and abs(c["rmax"] - 0.2 * side) < 1e-15
        assert c["rho_in"] == 1.0 and 10.0 <= c["rho_out"] <= 1000.0
    # every value of each draw shows up over 96 systems
    assert {c["nx"] for c in a} == dims and {c["nsph"] for c in a} == set(range(1, 9))
    # expected cells per system E[n]^3 = 73.6^3 ~ 4.0e5 (loose bound on 96 draws)
    mean = sum(c["nx"] * c["ny"] * c["nz"] for c in a) / len(a)
    assert 2.5e5 < mean < 6e5


def test_batched_abi_rejects_bad_arguments_without_a_device():
    L = psb.lib()
    b = C.c_void_p()
    assert L.psb_batch_create(None, 1, None) == -1
    assert L.psb_batch_create(C.byref(b), 0, None) == -1
    assert L.psb_batch_system(None, 0) is None
    assert L.psb_batch_destroy(None) == 0
    assert L.psb_batch_solve(None, None, None, None, None) == -1
    assert L.psb_batch_solve_device(None, None, None, None, None) == -1
    assert L.psb_batch_set_tuning(None, 0) == -1
    assert L.psb_batch_cases(C.c_ulonglong(1), -1, None) == -1


def test_oracle_level_rules_of_the_batched_and_semi_coarsened_hierarchies():
    # rule 2: halve a direction while it stays >= 4
    assert O.levels(2, 48, 32, 96) == [(48, 32, 96), (24, 16, 48), (12, 8, 24), (6, 4, 12),
                                       (6, 4, 6)]
    assert O.levels(2, 32, 32, 32)[-1] == (4, 4, 4)
    assert O.levels(2, 30, 8, 8) == [(30, 8, 8), (15, 4, 4)]
    # rule 1 stops at 8
    assert O.levels(1, 48, 32, 96)[-1] == (12, 8, 12)
    # rule 3: x and y first (while >= 8), then z
    assert O.levels(3, 32, 32, 64) == [(32, 32, 64), (16, 16, 64), (8, 8, 64), (8, 8, 32),
                                       (8, 8, 16), (8, 8, 8)]
    assert O.levels(3, 16, 48, 12) == [(16, 48, 12), (8, 24, 12), (8, 12, 12)]
