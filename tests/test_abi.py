"""libpsb.so loads and exports every entry point include/psb.h declares; the
calls that need no GPU behave as documented.  CPU only."""
import ctypes as C
import os
import re

import pytest

import psb

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "psb.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(psb_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_abi():
    syms = declared_symbols()
    assert "psb_solve" in syms and "psb_create" in syms
    assert set(syms) == set(psb.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = psb.lib()
    for s in declared_symbols():
        assert hasattr(L, s), s


def test_stop_defaults_are_the_reference_defaults():
    s = psb.stop_defaults()
    # MinIter 1, MaxIter 20, ResRat 0.01, ResTol 1e-15 (Src/Ravioli/*.h), scale 1
    assert (s.min_iter, s.max_iter, s.res_rat, s.res_tol, s.scale) == (1, 20, 0.01, 1e-15, 1.0)
    assert s.stalecount == -1 and s.precform == 1 and s.method == psb.PSB_CG
    assert s.precond == psb.PSB_PREC_DIAG and s.precision == psb.PSB_FP64


def test_structs_carry_their_size():
    # psb_stop_defaults / psb_cycle_defaults write sizeof() of the library's
    # structs: the ctypes mirrors must have the same layout, and a caller built
    # against another header version is recognised by the entry points
    s = psb.stop_defaults()
    assert s.struct_size == C.sizeof(psb.Stop)
    c = psb._cycle()
    assert c.struct_size == C.sizeof(psb.Cycle)
    assert s.ac_flexible == -1


def test_status_strings():
    L = psb.lib()
    assert L.psb_status_string(0) == b"converged"
    assert L.psb_status_string(-1) == b"invalid argument"
    assert L.psb_status_string(12345) == b"unknown status"


def test_invalid_arguments_are_rejected_without_a_device():
    L = psb.lib()
    ctx = C.c_void_p()
    assert L.psb_create(None, None) == -1
    assert L.psb_create(C.byref(ctx), C.byref(psb.Layout(0, 4, 4, 0))) == -1
    assert L.psb_create(C.byref(ctx), C.byref(psb.Layout(4, 4, 4, -1))) == -1
    assert L.psb_solve(None, None, None, None, None) == -1
    assert L.psb_destroy(None) == 0
    assert L.psb_generate(None, *([None] * 12)) == -1
