"""GPU parity of the FP32 bucketed container across many moves: particles cross brick
faces every step (the per-step re-bucketing with its new-bucket histogram in the gather
kernel and the counting-sort scatter), and clustered initial conditions go through
several PM steps without losing or duplicating a particle.

Every check is against the FP64 oracle on the container's own FP64 positions: cells and
per-cell counts bit-exact, drifted positions within one FP32 offset rounding, deposits
rel-L2 <= 1e-6 and accelerations rel-L2 <= 1e-5, as for a freshly sorted container.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import latfield2_amd as lf  # noqa: E402
import oracle as orc  # noqa: E402
from gpu_util import announce, real_np, rel_l2  # noqa: E402

pytestmark = pytest.mark.gpu


def _by_id(d):
    order = np.argsort(d["id"])
    return {k: v[order] for k, v in d.items()}


def _upload(N, pos, vel, mass=None):
    parts = lf.Particles(N, len(pos), lf.F32, ids=True, per_particle_mass=mass is not None)
    parts.addParticles(pos, vel, mass, ids=np.arange(len(pos)))
    return parts


def _check_state(parts, N, tag):
    """download, counts, the five deposits and a gather against the oracle"""
    d = _by_id(parts.download())
    pos = np.ascontiguousarray(d["pos"])
    assert np.array_equal(d["cell"], orc.bin_keys(pos, N)), tag
    assert np.array_equal(parts.cell_counts().cpu().numpy(), orc.cell_counts(pos, N)), tag
    lat = lf.Lattice(3, N, 1)
    for kind, C, proj, comm in ((lf.PROJ_SCALAR, 1, orc.scalar_project, orc.scalar_comm),
                                (lf.PROJ_VECTOR, 3, orc.vector_project, orc.vector_comm),
                                (lf.PROJ_VECVEC, 6, orc.vecvec_project, orc.vecvec_comm),
                                (lf.PROJ_VECTOR_NGP, 3, orc.vector_ngp_project, orc.vector_ngp_comm),
                                (lf.PROJ_SYMTENSOR_NGP, 6, orc.symtensor_ngp_project, orc.symtensor_ngp_comm)):
        f = lf.Field(lat, C, lf.F32)
        f.zero()
        parts.deposit(kind, f, lf.DEP_PERIODIC)
        torch.cuda.synchronize()
        ref = orc.real_field(N, 1, C)
        if kind == lf.PROJ_SCALAR:
            proj(pos, ref, N, 1, mass=d["mass"])
        else:
            proj(pos, d["vel"], ref, N, 1, mass=d["mass"])
        comm(ref, N, 1)
        s = slice(1, N + 1)
        err = rel_l2(real_np(f)[s, s, s], ref[s, s, s])
        assert err <= 1e-6, (tag, kind, err)
    return d


def _brick_edge(N):
    """the container's brick edge: the largest divisor of N up to 8"""
    return next(b for b in range(8, 0, -1) if N % b == 0)


def _bricks(cell, N, B=8):
    x, y, z = cell % N, (cell // N) % N, cell // (N * N)
    return x // B + (N // B) * (y // B + (N // B) * (z // B))


def _oracle_kick(d, phi, N, dt):
    phi_np = np.ascontiguousarray(real_np(phi))
    orc.ghost_fill(phi_np, N, 1)
    F = orc.force_field(phi_np, N, 1)
    v = np.ascontiguousarray(d["vel"].copy())
    orc.gather_kick(d["pos"], v, F, N, 1, dt_kick=dt)
    return v


@pytest.mark.parametrize("N,n", [(32, 40000),   # ~600 per brick: item path
                                 (32, 6000),    # ~90 per brick: chunks of sparse bricks
                                 (20, 6000),    # brick edge 5, 4^3 bricks
                                 (12, 1000)])   # brick edge 6, 2^3 bricks (periodic images coincide)
def test_steps_with_movers_match_oracle(N, n):
    """many brick crossings per step; every step checked on the positions it started from"""
    announce(f"layout movers N={N} n={n}")
    rng = np.random.default_rng(5)
    pos = rng.random((n, 3))
    vel = rng.standard_normal((n, 3)) * (1.5 / (N * 0.01))  # ~1.5 cells per step
    mass = rng.random(n) + 0.5
    parts = _upload(N, pos, vel, mass)
    lat = lf.Lattice(3, N, 1)
    phi = lf.Field(lat, 1, lf.F32)
    g = torch.Generator(device="cpu").manual_seed(3)
    phi.t.copy_(torch.randn(phi.t.shape, generator=g, dtype=torch.float64).to(torch.float32))
    moved = 0
    for step in range(12):
        d0 = _check_state(parts, N, f"step {step}")
        v_ref = _oracle_kick(d0, phi, N, 0.05)
        parts.gather_kick_drift(phi, dt_kick=0.05, dt_drift=0.01)
        torch.cuda.synchronize()
        d1 = _by_id(parts.download())
        assert rel_l2(d1["vel"] - d0["vel"], v_ref - d0["vel"]) <= 1e-5
        p = np.ascontiguousarray(d0["pos"])
        orc.drift(p, d1["vel"], 0.01)  # the drift uses the kicked FP32 velocity
        assert np.array_equal(d1["cell"], orc.bin_keys(p, N)), step
        diff = np.abs(d1["pos"] - p)
        assert np.max(np.minimum(diff, 1.0 - diff)) <= 2.0 ** -24 / N
        B = _brick_edge(N)
        moved += int(np.count_nonzero(_bricks(d1["cell"], N, B) != _bricks(d0["cell"], N, B)))
    print(f"brick crossings over 12 steps: {moved}")
    assert moved > n  # most particles crossed a brick face at least once
    _check_state(parts, N, "final")


def test_pm_steps_blob_ic_counts_conserved():
    """clustered IC through several PM steps: no particle lost or duplicated"""
    N = 64
    announce(f"layout pm steps blobs N={N}")
    np_ = N ** 3
    parts = lf.Particles(N, np_, lf.F32, ids=True, mass=1.0 / np_)
    parts.ic_blobs(np_, 64, 1.5, 2.0, 0.2, 0.05, 11)
    lat = lf.Lattice(3, N, 1)
    rho, phi = lf.Field(lat, 1, lf.F32), lf.Field(lat, 1, lf.F32)
    plan = lf.PlanFFT(rho, None)
    for _ in range(10):
        lf.pm_step(parts, plan, rho, phi, 0.05, 0.05)
    torch.cuda.synchronize()
    d = _by_id(parts.download())
    assert np.array_equal(d["id"], np.arange(np_))
    assert int(parts.cell_counts().sum().item()) == np_
    _check_state(parts, N, "after PM steps")
