"""Shared helpers of the GPU parity tests (the HIP path against the FP64 oracle)."""
from __future__ import annotations

import sys

import numpy as np


def announce(msg: str) -> None:
    """print the case before its GPU work starts, so a failing run names it"""
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def rel_l2(a, b) -> float:
    dt = np.complex128 if (np.iscomplexobj(a) or np.iscomplexobj(b)) else np.float64
    a = np.asarray(a, dtype=dt)
    b = np.asarray(b, dtype=dt)
    den = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (den if den > 0 else 1.0))


def to_np(t):
    return t.detach().cpu().numpy()


def kfield_np(kf):
    """device KField -> complex numpy [ky, kz, kx, c] with ghosts"""
    a = to_np(kf.t).astype(np.float64)
    return a[..., 0] + 1j * a[..., 1]


def real_np(f):
    return to_np(f.t).astype(np.float64)
