"""lf2mi_poisson_green_noalloc (the entry point that allocates nothing: its eigenvalue table is
a static device array) against lf2mi_poisson_green on the same k-fields, through the C ABI.
Two calls in a row, the second on a smaller lattice, so the second overwrites the first
call's table; plus the size refusal beyond the table."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

import latfield2_amd as lf
from latfield2_amd import _lib

from gpu_util import announce, kfield_np, rel_l2

pytestmark = pytest.mark.gpu

DT = {"f64": lf.F64, "f32": lf.F32}
# FP64: the table is computed by the device's sin instead of the host's (a last-bit difference
# at most); FP32: the products are rounded to FP32 either way
TOL = {"f64": 1e-14, "f32": 1e-7}


def _green(fn, k, res, scale):
    rc = getattr(_lib.load(), fn)(ctypes.c_void_p(k.t.data_ptr()), k.N, k.C, k.kh, float(res), float(scale), k.code, None)
    assert rc == 0, _lib.load().lf2mi_last_error()


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_green_noalloc_matches_green(prec):
    for N, kh, C in ((64, 1, 2), (24, 0, 1)):
        announce(f"poisson_green_noalloc {prec} N={N} kh={kh} C={C}")
        lat = lf.Lattice(3, N, 1)
        a = lf.KField(lat, kh, C, DT[prec])
        g = torch.Generator(device="cpu").manual_seed(N)
        a.t.copy_(torch.randn(a.t.shape, generator=g, dtype=torch.float64).to(a.t.dtype))
        b = lf.KField(lat, kh, C, DT[prec])
        b.t.copy_(a.t)
        _green("lf2mi_poisson_green", a, 1.0 / N, 1.0 / N ** 3)
        _green("lf2mi_poisson_green_noalloc", b, 1.0 / N, 1.0 / N ** 3)
        torch.cuda.synchronize()
        err = rel_l2(kfield_np(b), kfield_np(a))
        print(f"N={N} {prec}: rel-L2 {err:.2e}")
        assert err <= TOL[prec], err
        # the zero mode is 0 in both
        assert float(b.interior()[0, 0, 0].abs().max()) == 0.0


def test_green_noalloc_refuses_beyond_table():
    lat = lf.Lattice(3, 8, 1)
    k = lf.KField(lat, 0, 1, lf.F32)
    L = _lib.load()
    rc = L.lf2mi_poisson_green_noalloc(ctypes.c_void_p(k.t.data_ptr()), 2048, 1, 0, 1.0 / 2048, 1.0, lf.F32, None)
    assert rc != 0
    assert b"exceeds" in L.lf2mi_last_error()
