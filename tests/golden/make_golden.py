"""Writes the committed golden fixtures of tests/golden/*.npz.

The reference ships no fixture files (its KATs are programmatic and are restated
in tests/test_oracle.py), so these vectors come from numpy's pocketfft and the
independent formulation in numpy_pm.py.  Run from the repo root:

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import numpy_pm  # noqa: E402


def fft_fixture():
    rng = np.random.default_rng(20261202)
    N = 8
    r = rng.standard_normal((N, N, N))  # [z, y, x]
    K = np.fft.rfftn(r)                 # [kz, ky, kx], e^{-i}, unnormalised
    kin = rng.standard_normal((N, N, N // 2 + 1)) + 1j * rng.standard_normal((N, N, N // 2 + 1))
    # FFTW/pocketfft c2r semantics, unnormalised: N^3 * irfftn
    rout = np.fft.irfftn(kin, s=(N, N, N), axes=(0, 1, 2)) * N ** 3
    np.savez(os.path.join(HERE, "fft_n8.npz"), N=N, r=r, K=K, kin=kin, rout=rout)


def deposit_fixture():
    rng = np.random.default_rng(7)
    N = 4
    pos = rng.random((20, 3))
    pos[0] = [0.0, 0.0, 0.0]          # on a node
    pos[1] = [0.25, 0.5, 0.75]        # on a node
    pos[2] = [0.125, 0.375, 0.9375]   # link / plaquette centres
    pos[3] = [0.999, 0.001, 0.5]      # near the periodic faces
    vel = rng.standard_normal((20, 3))
    mass = rng.random(20) + 0.5
    rho, B, T = numpy_pm.deposits(pos, vel, mass, N)
    np.savez(os.path.join(HERE, "deposit_n4.npz"), N=N, pos=pos, vel=vel, mass=mass, rho=rho, B=B, T=T)


def deposit_ngp_fixture():
    rng = np.random.default_rng(8)
    N = 4
    pos = rng.random((20, 3))
    pos[0] = [0.0, 0.0, 0.0]          # on a node
    pos[1] = [0.25, 0.5, 0.75]        # on a node
    pos[2] = [0.125, 0.375, 0.9375]   # link / plaquette centres
    pos[3] = [0.999, 0.001, 0.5]      # near the periodic faces
    pos[4] = [0.2499, 0.7501, 0.0]    # either side of a cell face
    vel = rng.standard_normal((20, 3))
    mass = rng.random(20) + 0.5
    B, T = numpy_pm.deposits_ngp(pos, vel, mass, N)
    np.savez(os.path.join(HERE, "deposit_ngp_n4.npz"), N=N, pos=pos, vel=vel, mass=mass, B=B, T=T)


def pm_fixture():
    rng = np.random.default_rng(11)
    N = 8
    np_ = 64
    pos = rng.random((np_, 3))
    vel = 0.01 * rng.standard_normal((np_, 3))
    mass = np.ones(np_)
    out = numpy_pm.pm_step(pos, vel, mass, N, dt_kick=0.01, dt_drift=0.01)
    np.savez(os.path.join(HERE, "pm_n8.npz"), N=N, pos0=pos, vel0=vel, dt=0.01, **out)


if __name__ == "__main__":
    fft_fixture()
    deposit_fixture()
    deposit_ngp_fixture()
    pm_fixture()
    print("fixtures written to", HERE)
