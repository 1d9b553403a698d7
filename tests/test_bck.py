"""Checkpoints in the reference's .bck format (Scalar::save / Scalar::load,
scalar_save.cpp:20-32, scalar_load.cpp:29-39; name_file, global_name_file.cpp):
the library's writer and reader against the reference's own build (the
reference loads what the library wrote and the library loads what the
reference wrote, bit for bit; the file names agree).  CPU only; the
reference part runs where oracle/_ref/ref_driver has been built."""
import os
import re
import subprocess

import numpy as np
import pytest

import psb

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_driver")


def test_round_trip_and_layout(tmp_path):
    shape = (6, 5, 4)
    x = np.random.default_rng(1).standard_normal(int(np.prod(shape)))
    p = str(tmp_path / "a.bck")
    psb.save_bck(p, x, shape, bw=3)
    raw = open(p, "rb").read()
    assert np.frombuffer(raw[:12], dtype=np.int32).tolist() == [12, 11, 10]
    full = np.frombuffer(raw[12:], dtype=np.float64).reshape(12, 11, 10)
    assert np.array_equal(full[3:9, 3:8, 3:7].ravel(), x)
    assert np.array_equal(full[0, 3:8, 3:7], full[3, 3:8, 3:7])  # ghosts: nearest interior
    y, bw = psb.load_bck(p, shape)
    assert bw == 3 and np.array_equal(x, y)
    psb.save_bck(p, x, shape, bw=0)
    y, bw = psb.load_bck(p, shape)
    assert bw == 0 and np.array_equal(x, y)
    with pytest.raises(psb.PsbError):
        psb.load_bck(p, (6, 5, 5))


def test_file_names():
    assert psb.bck_name("press", 12, 0) == "press_p000_000012.bck"
    assert psb.bck_name("phi", -1, 3) == "phi_p003.bck"
    assert psb.bck_name("phi") == "phi.bck"


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="the reference build is not present")
def test_reference_reads_ours_and_we_read_the_references(tmp_path):
    shape = (6, 5, 4)
    x = np.random.default_rng(3).standard_normal(int(np.prod(shape)))
    ours = str(tmp_path / "ours.bck")
    psb.save_bck(ours, x, shape)
    out = subprocess.run([DRIVER, "bck", str(tmp_path), ours, *map(str, shape)], check=True,
                         capture_output=True, text=True).stdout
    assert np.array_equal(np.fromfile(str(tmp_path / "bck_loaded.f64")), x)
    y, bw = psb.load_bck(str(tmp_path / "ref.bck"), shape)
    assert bw == 3
    assert np.array_equal(y, np.fromfile(str(tmp_path / "bck_saved_interior.f64")))
    name = re.search(r"BCKNAME (\S+)", out).group(1)
    assert name == psb.bck_name("press", 12, 0)
