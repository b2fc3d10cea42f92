"""GPU parity of the whole PM step against the FP64 oracle (orc_pm_step).

The step: zero rho, periodic CIC deposit, mean-subtracted Poisson solve with
-1/K^2 and 1/N^3, CIC gather of -grad phi, kick, drift, re-bucket.  Inputs are
generated on the device (counter-based RNG), downloaded, and given to the oracle
as the exact FP64 positions the container represents.  Bars (north_star):
  * FP64 plain path: phi and velocities <= 1e-10 relative (summation order only);
  * FP32 path: phi, forces and velocity updates rel-L2 <= 1e-5 against FP64;
  * cell indices bit-exact except for particles whose FP64 position sits within
    1e-6 cells of a cell face (the FP32 kick moves them by ~1e-9 cells); those are
    counted and must all be boundary-ambiguous.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import latfield2_amd as lf  # noqa: E402
import oracle as orc  # noqa: E402
from gpu_util import announce, real_np, rel_l2  # noqa: E402

pytestmark = pytest.mark.gpu
DT = {"f32": lf.F32, "f64": lf.F64}


def _by_id(d):
    order = np.argsort(d["id"])
    return {k: v[order] for k, v in d.items()}


def _run_step(N, prec, ic, dt_kick, dt_drift, seed, mass):
    np_ = ic["np"] if "np" in ic else ic["ng"] ** 3
    parts = lf.Particles(N, np_, DT[prec], ids=True, mass=mass)
    if ic["kind"] == "uniform":
        parts.ic_uniform(np_, ic.get("sigma_v", 0.0), seed)
    elif ic["kind"] == "lattice":
        parts.ic_lattice(ic["ng"], ic["sigma_cells"], ic["sigma_v"], seed)
    else:
        parts.ic_blobs(np_, ic["nblobs"], 1.5, 2.0, 0.2, ic.get("sigma_v", 0.0), seed)
    d0 = _by_id(parts.download())
    # binning of the initial state: cell keys and per-cell counts bit-exact
    assert np.array_equal(d0["cell"], orc.bin_keys(d0["pos"], N))
    assert np.array_equal(parts.cell_counts().cpu().numpy(), orc.cell_counts(d0["pos"], N))
    lat = lf.Lattice(3, N, 1)
    rho = lf.Field(lat, 1, DT[prec])
    phi = lf.Field(lat, 1, DT[prec])
    plan = lf.PlanFFT(rho, None)
    lf.pm_step(parts, plan, rho, phi, dt_kick, dt_drift)
    torch.cuda.synchronize()
    d1 = _by_id(parts.download())
    return d0, d1, real_np(rho), real_np(phi)


def _check_step(N, prec, d0, d1, rho_g, phi_g, dt_kick, dt_drift, mass, tol):
    pos = np.ascontiguousarray(d0["pos"].copy())
    vel = np.ascontiguousarray(d0["vel"].copy())
    out = orc.pm_step(pos, vel, N, dt_kick, dt_drift, mass_global=mass)
    s = slice(1, N + 1)
    rho_o = out["rho"][s, s, s, 0]            # mean subtracted
    rg = rho_g[s, s, s, 0]
    assert rel_l2(rg - rg.mean(), rho_o) <= tol
    assert rel_l2(phi_g[s, s, s, 0], out["phi"][s, s, s, 0]) <= tol
    assert rel_l2(d1["vel"], vel) <= tol
    keys = orc.bin_keys(pos, N)
    mism = np.nonzero(d1["cell"] != keys)[0]
    t = pos[mism] * N
    ambiguous = np.min(np.abs(t - np.round(t)), axis=1) < 1e-6
    assert np.all(ambiguous), f"{len(mism)} cell mismatches, {int(np.sum(~ambiguous))} not at a face"
    diff = np.abs(d1["pos"] - pos)
    diff = np.minimum(diff, 1.0 - diff)
    assert np.max(diff) * N <= (1e-9 if prec == "f64" else 1e-5)
    return out


@pytest.mark.parametrize("prec,tol", [("f64", 1e-10), ("f32", 1e-5)])
def test_pm_step_c1_uniform_forces(prec, tol):
    """64^3 lattice, 32^3 uniform particles of mass 1 at rest: rho, phi and the gathered
    force (read as the velocity after a unit kick with no drift)"""
    N = 64
    announce(f"pm_step C1 {prec}")
    d0, d1, rho, phi = _run_step(N, prec, {"kind": "uniform", "np": 32 ** 3}, 1.0, 0.0, seed=1, mass=1.0)
    out = _check_step(N, prec, d0, d1, rho, phi, 1.0, 0.0, 1.0, tol)
    assert np.array_equal(d1["cell"], d0["cell"])
    assert rel_l2(d1["vel"], out["acc"]) <= tol


@pytest.mark.parametrize("prec,tol", [("f64", 1e-10), ("f32", 1e-5)])
@pytest.mark.parametrize("N", [64, 96, 256])
def test_pm_step_lattice_gaussian(N, prec, tol):
    """the kick-drift configuration (256^3 mesh with 256^3 particles is its full size): grid +
    N(0, 0.3 cells), v ~ N(0, 0.01), dt 0.01, mass 1/Np (mean density 1)"""
    dt = 0.01
    announce(f"pm_step lattice {N} {prec}")
    ic = {"kind": "lattice", "ng": N, "sigma_cells": 0.3, "sigma_v": 0.01}
    d0, d1, rho, phi = _run_step(N, prec, ic, dt, dt, seed=2, mass=1.0 / N ** 3)
    _check_step(N, prec, d0, d1, rho, phi, dt, dt, 1.0 / N ** 3, tol)


@pytest.mark.parametrize("prec,tol", [("f32", 1e-5), ("f64", 1e-10)])
def test_pm_step_blobs_64(prec, tol):
    """clustered Pareto blobs (sigma 2 cells, 20% uniform) at 64^3, mass 1/Np"""
    N, dt = 64, 0.01
    announce(f"pm_step blobs 64 {prec}")
    ic = {"kind": "blobs", "np": N ** 3, "nblobs": 512, "sigma_v": 0.01}
    d0, d1, rho, phi = _run_step(N, prec, ic, dt, dt, seed=3, mass=1.0 / N ** 3)
    _check_step(N, prec, d0, d1, rho, phi, dt, dt, 1.0 / N ** 3, tol)


def test_clustered_deposit_512_float_and_fixed_point():
    """512^3 mesh, 512^3 particles in 4096 Pareto(1.5) blobs (sigma 2 cells, 20% uniform):
    the FP32 deposit with float atomics and with int64 fixed point (2^-32) against the FP64
    oracle deposit of the same positions; the fixed-point deposit is bitwise reproducible;
    then a kick-only PM step from the fixed-point deposit against the oracle step"""
    N = 512
    np_ = N ** 3
    mass = 1.0 / np_
    announce("clustered deposit 512^3, float and int64 fixed point")
    parts = lf.Particles(N, np_, lf.F32, ids=True, mass=mass)
    parts.ic_blobs(np_, 4096, 1.5, 2.0, 0.2, 0.01, 3)
    lat = lf.Lattice(3, N, 1)
    s = slice(1, N + 1)
    got = {}
    first_fixed = None
    for name, acc in (("native", lf.ACC_NATIVE), ("fixed", lf.ACC_FIXED64), ("fixed again", lf.ACC_FIXED64)):
        f = lf.Field(lat, 1, lf.F32)
        f.zero()
        parts.deposit(lf.PROJ_SCALAR, f, lf.DEP_PERIODIC, acc)
        torch.cuda.synchronize()
        if name == "fixed again":
            assert torch.equal(f.t, first_fixed)  # run-to-run reproducible
        else:
            got[name] = f.t[s, s, s, 0].double().cpu().numpy()
            if name == "fixed":
                first_fixed = f.t.clone()
        del f
    del first_fixed
    d0 = _by_id(parts.download())
    orc.set_threads(16)
    ref = orc.real_field(N, 1, 1)
    orc.scalar_project(d0["pos"], ref, N, 1, mass_global=mass)
    orc.scalar_comm(ref, N, 1)
    errs = {k: rel_l2(v, ref[s, s, s, 0]) for k, v in got.items()}
    print(f"deposit rel-L2 against the FP64 oracle: {errs}")
    assert errs["native"] <= 1e-6 and errs["fixed"] <= 1e-6
    del ref, got
    # kick-only step (no drift) from the fixed-point deposit
    rho = lf.Field(lat, 1, lf.F32)
    phi = lf.Field(lat, 1, lf.F32)
    plan = lf.PlanFFT(rho, None)
    lf.pm_step(parts, plan, rho, phi, 0.01, 0.0, acc=lf.ACC_FIXED64)
    torch.cuda.synchronize()
    d1 = _by_id(parts.download())
    pos = np.ascontiguousarray(d0["pos"].copy())
    vel = np.ascontiguousarray(d0["vel"].copy())
    out = orc.pm_step(pos, vel, N, 0.01, 0.0, mass_global=mass, want_acc=False)
    e_phi = rel_l2(real_np(phi)[s, s, s, 0], out["phi"][s, s, s, 0])
    e_dv = rel_l2(d1["vel"] - d0["vel"], vel - d0["vel"])
    print(f"phi rel-L2 {e_phi:.2e}, velocity-change rel-L2 {e_dv:.2e}")
    assert np.array_equal(d1["cell"], d0["cell"])
    assert e_phi <= 1e-5 and e_dv <= 1e-5


def test_pm_step_replays_as_a_hip_graph():
    """A PM step (zero, fixed-point deposit, Poisson, gather + kick + drift, re-bucket) issues
    only kernels on the caller's stream, so steps can be captured once and replayed.  The
    container alternates two particle buffers, so a replayable capture holds an even number of
    steps (here 2).  Replays on one container match eager steps on an identical one bit for bit
    (the fixed-point deposit makes the fields independent of the particle order), with eager
    work and synchronisations between the replays."""
    N, Np = 64, 64 ** 3
    announce(f"pm step as a HIP graph N={N}")
    lat = lf.Lattice(3, N, 1)

    def make():
        parts = lf.Particles(N, Np, lf.F32, mass=1.0 / Np)
        parts.ic_blobs(Np, nblobs=64, alpha=1.5, sigma_cells=2.0, frac_uniform=0.2, sigma_v=0.01, seed=9)
        rho, phi = lf.Field(lat, 1, lf.F32), lf.Field(lat, 1, lf.F32)
        return parts, rho, phi, lf.PlanFFT(rho, None)

    a, b = make(), make()
    dt = 0.01
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        for st in (a, b):  # eager warmup: every buffer of the step exists before the capture
            lf.pm_step(st[0], st[3], st[1], st[2], dt, dt, lf.ACC_FIXED64)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        for _ in range(2):
            lf.pm_step(b[0], b[3], b[1], b[2], dt, dt, lf.ACC_FIXED64)
    # the capture did not run the steps: each replay is matched by two eager steps on a
    for it in range(3):
        for _ in range(2):
            lf.pm_step(a[0], a[3], a[1], a[2], dt, dt, lf.ACC_FIXED64)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(a[1].t, b[1].t), it   # rho
        assert torch.equal(a[2].t, b[2].t), it   # phi
        assert torch.equal(a[0].cell_counts(), b[0].cell_counts()), it
