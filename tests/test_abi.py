"""CPU checks of the C ABI boundary: the library loads, exports every entry point that
include/lf2mi.h declares, and rejects bad arguments on the host (no GPU work)."""
import ctypes
import os
import re

import pytest

import latfield2_amd as lf
from latfield2_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lf2mi.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lf2mi_\w+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ["lf2mi_ghost_fill", "lf2mi_projection_comm", "lf2mi_fft_plan_create", "lf2mi_fft_execute",
                 "lf2mi_poisson_green", "lf2mi_poisson_solve", "lf2mi_deposit", "lf2mi_gather_kick_drift",
                 "lf2mi_pm_step", "lf2mi_last_error"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    L = lf.load()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    assert set(declared_functions()) == set(_lib.SIGNATURES)


@pytest.mark.parametrize("call,args,code", [
    ("lf2mi_fft_plan_create", (14, 1, 1, 0, 0), -4),      # N not 2^k, 3 * 2^k or 5 * 2^k
    ("lf2mi_fft_plan_create", (36, 1, 1, 0, 0), -4),      # 9 * 2^2: no radix-9 plan
    ("lf2mi_fft_plan_create", (6, 1, 1, 0, 0), -4),       # below the supported range
    ("lf2mi_fft_plan_create", (2048, 1, 1, 0, 0), -4),    # above the supported range
    ("lf2mi_fft_plan_create", (64, 0, 1, 0, 0), -1),      # no components
    ("lf2mi_fft_plan_create", (64, 1, 1, 0, 7), -1),      # bad precision
])
def test_plan_create_rejects_bad_geometry(call, args, code):
    h = ctypes.c_void_p()
    rc = getattr(lf.load(), call)(*args, ctypes.byref(h))
    assert rc == code and not h.value
    assert lf.load().lf2mi_last_error().decode()


def test_host_side_argument_checks():
    L = lf.load()
    assert L.lf2mi_zero(None, -1, 0, None) == -1
    assert L.lf2mi_zero(None, 10, 0, None) == -1
    # halo 0 is rejected by the folds exactly as the reference exits (projections.hpp:79-84)
    assert L.lf2mi_projection_comm(ctypes.c_void_p(1), 8, 0, 0, 0, None) == -1
    assert b"halo" in L.lf2mi_last_error()
    assert L.lf2mi_projection_comm(ctypes.c_void_p(1), 8, 1, 9, 0, None) == -1
    assert L.lf2mi_ghost_fill(None, 8, 1, 1, 0, None) == -1
    assert L.lf2mi_ghost_fill(ctypes.c_void_p(1), 8, 9, 1, 0, None) == -1  # halo larger than N
    assert L.lf2mi_poisson_green(None, 8, 1, 0, 0.1, 1.0, 0, None) == -1
    assert L.lf2mi_poisson_green_noalloc(None, 8, 1, 0, 0.1, 1.0, 0, None) == -1
    # the no-allocation Green's multiply has a 1024-entry table: larger N is refused before any
    # device work
    assert L.lf2mi_poisson_green_noalloc(ctypes.c_void_p(1), 2048, 1, 0, 0.1, 1.0, 0, None) == -4
    assert b"exceeds" in L.lf2mi_last_error()
    assert L.lf2mi_deposit(None, 0, None, 1, 0, 0, None) == -1
    h = ctypes.c_void_p()
    assert L.lf2mi_parts_create(1, 10, 0, 0, ctypes.byref(h)) == -4
    assert L.lf2mi_parts_create(64, -5, 0, 0, ctypes.byref(h)) == -4
    assert L.lf2mi_parts_create(64, 10, 3, 0, ctypes.byref(h)) == -1
    # unknown FFT directions are a no-op in the reference; a null plan is an error here
    assert L.lf2mi_fft_execute(None, 1, None, None, None) == -1
