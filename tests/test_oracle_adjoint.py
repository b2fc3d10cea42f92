"""Pins the FP64 adjoint of the multi-step PM (oracle/lf2_oracle.c orc_diffpm_*) against
central finite differences of the FP64 forward, and the loss against an independent
numpy evaluation of the same forward (tests/golden/numpy_pm.py pieces).  A dropped
term of the adjoint (the direct interpolation term, the field term through the Poisson
solve, or the drift's velocity coupling) changes the gradient by far more than the
tolerance."""
import numpy as np
import pytest

import oracle as orc


def _case(N=16, n=16, sigma=0.5, seed=0):
    rng = np.random.default_rng(seed)
    g = (np.arange(n) + 0.5) / n
    base = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    pos = (base + sigma / N * rng.standard_normal(base.shape)) % 1.0
    vel = 0.01 * rng.standard_normal(base.shape)
    tgt_pos = (base + sigma / N * np.random.default_rng(seed + 1000).standard_normal(base.shape)) % 1.0
    return pos, vel, tgt_pos


def test_forward_density_matches_numpy_formulation():
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    import numpy_pm
    pos, vel, _ = _case()
    N, mass = 16, 1.0 / 4096
    rho = orc.cic_density(pos, N, mass)
    assert np.allclose(rho, numpy_pm.hat_deposit(pos, N, np.full(len(pos), mass * N ** 3)), rtol=1e-12)
    # one step against the numpy PM step (same kick-drift, same Poisson operator)
    p1, v1 = orc.diffpm_forward(pos, vel, N, mass, 0.02, 1)
    ref = numpy_pm.pm_step(pos, vel, np.full(len(pos), mass), N, 0.02, 0.02)
    assert np.allclose(v1, ref["vel"], rtol=1e-10, atol=1e-14)
    assert np.allclose(p1, ref["pos"], rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("nsteps", [1, 3])
def test_adjoint_matches_central_differences(nsteps):
    N, mass, dt = 16, 1.0 / 4096, 0.02
    pos, vel, tpos = _case()
    target = orc.cic_density(tpos, N, mass)
    L, gp, gv = orc.diffpm_loss_grad(pos, vel, N, mass, dt, nsteps, target)
    assert L > 0
    rng = np.random.default_rng(5)
    idx = rng.choice(pos.size, 24, replace=False)
    h = 1e-3 / N  # 1e-3 cells
    checked = 0
    for i in idx:
        def loss_at(dx):
            p = pos.copy().reshape(-1)
            p[i] += dx
            return orc.diffpm_loss(p.reshape(-1, 3), vel, N, mass, dt, nsteps, target)
        fd = (loss_at(h) - loss_at(-h)) / (2 * h)
        fd2 = (loss_at(h / 2) - loss_at(-h / 2)) / h
        if abs(fd - fd2) > 1e-6 * max(abs(fd), 1e-30) + 1e-12 * abs(L) / h:
            continue  # the perturbation crosses a CIC kink: FD is not a derivative there
        checked += 1
        assert abs(gp.reshape(-1)[i] - fd) <= 1e-5 * max(abs(fd), 1e-3 * np.abs(gp).max()), (i, gp.reshape(-1)[i], fd)
    assert checked >= 18
    # velocity gradient of a one-coordinate perturbation as well
    j = int(idx[0])
    def lv(dv):
        v = vel.copy().reshape(-1)
        v[j] += dv
        return orc.diffpm_loss(pos, v.reshape(-1, 3), N, mass, dt, nsteps, target)
    hv = 1e-6
    fdv = (lv(hv) - lv(-hv)) / (2 * hv)
    assert abs(gv.reshape(-1)[j] - fdv) <= 1e-5 * max(abs(fdv), 1e-3 * np.abs(gv).max())
