"""Differentiable PM on the GPU: the hand-written adjoint kernels of lf2mi_diffpm_*
against the FP64 CPU adjoint of the oracle, against central finite differences of the
FP64 GPU forward, and through the torch.autograd.Function harness.

Workload of the full case: 128^3 mesh, 128^3 particles on the lattice plus a seeded
Gaussian displacement (sigma 0.5 cells), zero initial velocity, 10 leapfrog steps at
dt = 0.02, mass 1/Np (mean density 1).  The target is the final density of the same run
from an independent seed.  Finite differences use h = 1e-3 cells on 64 random
coordinates of particles whose FP64 trajectory keeps at least 10 h from every cell face
along the perturbed axis (the CIC loss has kinks at cell faces, where a central
difference does not measure the derivative).
"""
from __future__ import annotations

import numpy as np
import pytest

from gpu_util import announce, rel_l2

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _lib():
    import latfield2_amd as lf
    from latfield2_amd.diffpm import DiffPM, pm_density_loss
    lf.load()
    return lf, DiffPM, pm_density_loss


def lattice_positions(N: int, ng: int, sigma_cells: float, seed: int) -> np.ndarray:
    """particles on an ng^3 grid (index x + ng(y + ng z)) plus a Gaussian displacement, box units"""
    rng = np.random.default_rng(seed)
    iz, iy, ix = np.meshgrid(np.arange(ng), np.arange(ng), np.arange(ng), indexing="ij")
    grid = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=1).astype(np.float64) * (N / ng)
    return np.mod((grid + rng.normal(0.0, sigma_cells, grid.shape)) / N, 1.0)


def dev(a, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device="cuda", dtype=dtype).contiguous()


@pytest.fixture(scope="module")
def small():
    """16^3 mesh, 16^3 particles, 3 steps: every GPU precision against the oracle"""
    N, nsteps, dt = 16, 3, 0.02
    np_ = N ** 3
    mass = 1.0 / np_
    pos0 = lattice_positions(N, N, 0.5, seed=3)
    rng = np.random.default_rng(4)
    vel0 = rng.normal(0.0, 0.05, pos0.shape)
    import oracle as orc
    tpos = lattice_positions(N, N, 0.5, seed=1003)
    tp, _ = orc.diffpm_forward(tpos, np.zeros_like(tpos), N, mass, dt, nsteps)
    target = orc.cic_density(tp, N, mass)
    L, gp, gv = orc.diffpm_loss_grad(pos0, vel0, N, mass, dt, nsteps, target)
    fpos, fvel = orc.diffpm_forward(pos0, vel0, N, mass, dt, nsteps)
    return dict(N=N, nsteps=nsteps, dt=dt, mass=mass, pos0=pos0, vel0=vel0, target=target,
                L=L, gp=gp, gv=gv, fpos=fpos, fvel=fvel)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_diffpm_small_matches_oracle(small, prec):
    lf, DiffPM, _ = _lib()
    s = small
    announce(f"diffpm small N={s['N']} steps={s['nsteps']} {prec}")
    code = lf.F64 if prec == "f64" else lf.F32
    m = DiffPM(s["N"], s["N"] ** 3, s["nsteps"], s["mass"], s["dt"], code)
    L, rho = m.forward(dev(s["pos0"]), dev(s["vel0"]), dev(s["target"].ravel()), want_rho=True)
    gp, gv = m.backward()
    torch.cuda.synchronize()
    pos, vel = m.state(s["nsteps"])
    dpos = np.mod(pos.cpu().numpy() - s["fpos"] + 0.5, 1.0) - 0.5
    err_L = abs(L - s["L"]) / s["L"]
    err_gp = rel_l2(gp.cpu().numpy(), s["gp"])
    err_gv = rel_l2(gv.cpu().numpy(), s["gv"])
    print(f"loss {L:.10e} vs {s['L']:.10e}: {err_L:.2e}; pos max {np.abs(dpos).max() * s['N']:.2e} cells; "
          f"grad pos {err_gp:.2e}, grad vel {err_gv:.2e}")
    if prec == "f64":
        assert err_L <= 1e-11
        assert np.abs(dpos).max() * s["N"] <= 1e-10
        assert rel_l2(vel.cpu().numpy(), s["fvel"]) <= 1e-11
        assert err_gp <= 1e-9 and err_gv <= 1e-9
    else:
        assert err_L <= 1e-5
        assert np.abs(dpos).max() * s["N"] <= 1e-4
        assert err_gp <= 1e-4 and err_gv <= 1e-4
    # the final density is the periodic CIC density of the final positions
    ref_rho = __import__("oracle").cic_density(s["fpos"], s["N"], s["mass"])
    assert rel_l2(rho.cpu().numpy(), ref_rho) <= (1e-12 if prec == "f64" else 1e-5)


@pytest.fixture(scope="module")
def c4():
    lf, DiffPM, _ = _lib()
    N, nsteps, dt = 128, 10, 0.02
    np_ = N ** 3
    mass = 1.0 / np_
    announce(f"diffpm C4 setup N={N} np={np_} steps={nsteps}")
    pos0 = lattice_positions(N, N, 0.5, seed=1)
    tpos = lattice_positions(N, N, 0.5, seed=1001)
    m64 = DiffPM(N, np_, nsteps, mass, dt, lf.F64)
    _, target = m64.forward(dev(tpos), None, None, want_rho=True)
    target = target.reshape(-1).contiguous()
    p0 = dev(pos0)
    L64 = m64.forward(p0, None, target)
    g64, _ = m64.backward(want_vel=False)
    # closest approach to a cell face along each axis over the FP64 trajectory
    dmin = torch.full((np_, 3), 1.0, dtype=torch.float64, device="cuda")
    for s in range(nsteps + 1):
        ps, _ = m64.state(s)
        f = ps * N - torch.floor(ps * N)
        dmin = torch.minimum(dmin, torch.minimum(f, 1.0 - f))
    m32 = DiffPM(N, np_, nsteps, mass, dt, lf.F32)
    L32 = m32.forward(p0, None, target)
    g32, _ = m32.backward(want_vel=False)
    torch.cuda.synchronize()
    return dict(N=N, np=np_, nsteps=nsteps, dt=dt, mass=mass, pos0=pos0, p0=p0, target=target, m64=m64,
                L64=L64, g64=g64, L32=L32, g32=g32, dmin=dmin)


def test_diffpm_c4_adjoint_vs_finite_differences(c4):
    h = 1e-3  # cells
    N = c4["N"]
    rng = np.random.default_rng(2024)
    dmin = c4["dmin"].cpu().numpy()
    picks = []
    while len(picks) < 64:
        p, d = int(rng.integers(c4["np"])), int(rng.integers(3))
        if dmin[p, d] > 10 * h and (p, d) not in picks:
            picks.append((p, d))
    announce(f"diffpm C4 finite differences on {len(picks)} coordinates, h={h} cells")
    m64, p0, tgt = c4["m64"], c4["p0"], c4["target"]
    fd = np.zeros(len(picks))
    for i, (p, d) in enumerate(picks):
        lp = m64.forward(p0.index_put((torch.tensor([p], device="cuda"), torch.tensor([d], device="cuda")),
                                      p0[p, d] + h / N), None, tgt)
        lm = m64.forward(p0.index_put((torch.tensor([p], device="cuda"), torch.tensor([d], device="cuda")),
                                      p0[p, d] - h / N), None, tgt)
        fd[i] = (lp - lm) / (2 * h)  # dL/dX per cell
    idx = tuple(np.array(picks).T)
    a64 = c4["g64"].cpu().numpy()[idx] / N
    a32 = c4["g32"].cpu().numpy()[idx] / N
    e64, e32 = rel_l2(a64, fd), rel_l2(a32, fd)
    worst = np.max(np.abs(a32 - fd) / np.maximum(np.abs(fd), 1e-12 * np.abs(fd).max()))
    print(f"FD |g| {np.linalg.norm(fd):.4e}: FP64 adjoint rel-L2 {e64:.2e}, FP32 adjoint rel-L2 {e32:.2e}, "
          f"FP32 worst coordinate {worst:.2e}")
    assert e64 <= 1e-6
    assert e32 <= 1e-4


def test_diffpm_c4_against_fp64_cpu_adjoint(c4):
    import oracle as orc
    announce("diffpm C4 FP64 CPU adjoint")
    orc.set_threads(16)
    L, gp, _ = orc.diffpm_loss_grad(c4["pos0"], np.zeros_like(c4["pos0"]), c4["N"], c4["mass"], c4["dt"],
                                    c4["nsteps"], c4["target"].cpu().numpy())
    e64 = rel_l2(c4["g64"].cpu().numpy(), gp)
    e32 = rel_l2(c4["g32"].cpu().numpy(), gp)
    print(f"loss cpu {L:.10e} gpu64 {c4['L64']:.10e} gpu32 {c4['L32']:.10e}; "
          f"grad rel-L2 FP64 {e64:.2e} FP32 {e32:.2e}")
    assert abs(c4["L64"] - L) / L <= 1e-10
    assert abs(c4["L32"] - L) / L <= 1e-5
    assert e64 <= 1e-8
    assert e32 <= 1e-4


def test_diffpm_checkpoints_fit_in_hbm(c4):
    b = c4["m64"].bytes()
    # 11 FP64 states of 52 B per particle plus fields and adjoints
    assert b < 4 * 2 ** 30, b
    print(f"FP64 C4 handle holds {b / 2 ** 20:.1f} MiB")


def test_autograd_harness():
    lf, DiffPM, pm_density_loss = _lib()
    N, nsteps, dt = 16, 2, 0.02
    np_ = N ** 3
    announce("diffpm autograd harness")
    pos0 = dev(lattice_positions(N, N, 0.5, seed=5), torch.float32).requires_grad_()
    vel0 = torch.zeros_like(pos0).requires_grad_()
    target = (torch.rand(N ** 3, device="cuda", dtype=torch.float32, generator=None) * 2).requires_grad_()
    m = DiffPM(N, np_, nsteps, 1.0 / np_, dt, lf.F32)
    loss = pm_density_loss(pos0, vel0, target, m)
    (3.0 * loss).backward()
    m2 = DiffPM(N, np_, nsteps, 1.0 / np_, dt, lf.F32)
    L2, rho = m2.forward(pos0.detach(), vel0.detach(), target.detach(), want_rho=True)
    gp, gv = m2.backward()
    torch.cuda.synchronize()
    assert abs(loss.item() - L2) <= 1e-6 * L2
    assert rel_l2(pos0.grad.cpu().numpy(), 3.0 * gp.cpu().numpy()) <= 1e-6
    assert rel_l2(vel0.grad.cpu().numpy(), 3.0 * gv.cpu().numpy()) <= 1e-6
    gt = -6.0 * (rho.reshape(-1) - target.detach())
    assert rel_l2(target.grad.cpu().numpy(), gt.cpu().numpy()) <= 1e-6


def test_autograd_two_evaluations_of_one_model_refuse_the_stale_backward():
    """L(a) + L(b) on one DiffPM: the device holds the checkpoints of b only, so the first
    node's backward must raise rather than return b's gradients for a"""
    lf, DiffPM, pm_density_loss = _lib()
    N, nsteps, dt = 16, 1, 0.02
    np_ = N ** 3
    announce("diffpm two evaluations before one backward")
    m = DiffPM(N, np_, nsteps, 1.0 / np_, dt, lf.F32)
    a = dev(lattice_positions(N, N, 0.5, seed=6), torch.float32).requires_grad_()
    b = dev(lattice_positions(N, N, 0.5, seed=7), torch.float32).requires_grad_()
    loss = pm_density_loss(a, None, None, m) + pm_density_loss(b, None, None, m)
    with pytest.raises(RuntimeError, match="checkpoints belong to forward"):
        loss.backward()
    # one model per evaluation works
    m2 = DiffPM(N, np_, nsteps, 1.0 / np_, dt, lf.F32)
    a.grad = None
    b.grad = None
    loss = pm_density_loss(a, None, None, m) + pm_density_loss(b, None, None, m2)
    loss.backward()
    ma = DiffPM(N, np_, nsteps, 1.0 / np_, dt, lf.F32)
    ma.forward(a.detach())
    ga, _ = ma.backward(want_vel=False)
    torch.cuda.synchronize()
    assert rel_l2(a.grad.cpu().numpy(), ga.cpu().numpy()) <= 1e-6


def test_diffpm_errors():
    lf, DiffPM, _ = _lib()
    from latfield2_amd import _lib as raw
    m = DiffPM(16, 64, 1, 1.0, 0.01, lf.F32)
    g = torch.empty((64, 3), dtype=torch.float32, device="cuda")
    with pytest.raises(lf.Lf2miError):  # no forward pass to differentiate
        raw.call("lf2mi_diffpm_backward", m.h, g.data_ptr(), None, None)
    with pytest.raises(lf.Lf2miError):
        DiffPM(12, 64, 1, 1.0, 0.01, lf.F32)  # N must be a power of two
