"""GPU parity of the boundary pieces a driver calls around the PM kernels:

* lf2mi_field_sum (the driver's parallel.sum over a field, examples/poissonSolver.cpp:88-94)
  against numpy's FP64 sum of the same field, and of the oracle's deposit (mass
  conservation); lf2mi_field_affine against numpy;
* Particles::updateVel / moveParticles with the reference's signature
  (particles/LATfield2_Particles.hpp:217-244): the returned sqrt(max |v|^2) and the
  output[] slots reduced by SUM / MIN / MAX (:788-807, :846-885) against the same
  statistics of the oracle's kicked velocities;
* the scalar fixed-point deposit with per-particle masses of order 1/Np.

Bars: sums and affine maps of FP64 fields <= 1e-13 relative, of FP32 fields <= 1e-12 of the
FP64 sum of the FP32 values; reductions of the FP64 path <= 1e-12, of the FP32 path <= 1e-5
(the FP32 velocities of the kick); deposits <= 1e-6 rel-L2 against the FP64 oracle.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import latfield2_amd as lf  # noqa: E402
import oracle as orc  # noqa: E402
from gpu_util import announce, real_np, rel_l2  # noqa: E402

pytestmark = pytest.mark.gpu
DT = {"f32": lf.F32, "f64": lf.F64}


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("N,h,C,comp", [(16, 1, 1, 0), (20, 2, 3, 2), (64, 1, 6, 4)])
def test_field_sum_matches_numpy(prec, N, h, C, comp):
    announce(f"field_sum {prec} N={N} h={h} C={C} comp={comp}")
    f = lf.Field(lf.Lattice(3, N, h), C, DT[prec])
    g = torch.Generator(device="cpu").manual_seed(N + C)
    f.t.copy_((torch.randn(f.t.shape, generator=g, dtype=torch.float64) + 0.5).to(f.t.dtype))
    got = f.sum(comp)
    a = real_np(f)
    want = float(np.sum(a[h:N + h, h:N + h, h:N + h, comp]))
    scale = float(np.sum(np.abs(a[h:N + h, h:N + h, h:N + h, comp])))
    assert abs(got - want) <= (1e-13 if prec == "f64" else 1e-12) * scale


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_field_sum_of_a_deposit_is_the_oracle_mass(prec):
    """sum of the periodic deposit = sum of the oracle's folded deposit = N^3 sum m"""
    N, n = 32, 20000
    announce(f"field_sum of a deposit {prec}")
    rng = np.random.default_rng(3)
    pos, vel, mass = rng.random((n, 3)), rng.standard_normal((n, 3)), rng.random(n) + 0.5
    parts = lf.Particles(N, n, DT[prec], ids=True, per_particle_mass=True)
    parts.addParticles(pos, vel, mass, ids=np.arange(n))
    d = parts.download()
    order = np.argsort(d["id"])
    lat = lf.Lattice(3, N, 1)
    rho = lf.Field(lat, 1, DT[prec])
    lf.projection_init(rho)
    parts.deposit(lf.PROJ_SCALAR, rho, lf.DEP_PERIODIC)
    got = rho.sum(0)
    ref = orc.real_field(N, 1, 1)
    orc.scalar_project(np.ascontiguousarray(d["pos"][order]), ref, N, 1, mass=np.ascontiguousarray(d["mass"][order]))
    orc.scalar_comm(ref, N, 1)
    want = float(np.sum(ref[1:N + 1, 1:N + 1, 1:N + 1, 0]))
    assert abs(want / (N ** 3 * np.sum(d["mass"])) - 1) <= 1e-12
    assert abs(got - want) <= (1e-12 if prec == "f64" else 1e-6) * abs(want)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_field_affine_matches_numpy(prec):
    N, h, C, comp = 24, 2, 3, 1
    announce(f"field_affine {prec}")
    f = lf.Field(lf.Lattice(3, N, h), C, DT[prec])
    g = torch.Generator(device="cpu").manual_seed(11)
    f.t.copy_(torch.randn(f.t.shape, generator=g, dtype=torch.float64).to(f.t.dtype))
    a = real_np(f).copy()
    mul, add = 1.0 / N ** 3, -0.375
    f.affine(mul, add, comp)
    got = real_np(f)
    want = a.copy()
    want[h:N + h, h:N + h, h:N + h, comp] = a[h:N + h, h:N + h, h:N + h, comp] * mul + add
    if prec == "f32":
        want = want.astype(np.float32).astype(np.float64)
        ulp = np.spacing(np.abs(want).astype(np.float32)).astype(np.float64)
        assert np.all(np.abs(got - want) <= ulp)
    else:
        assert np.max(np.abs(got - want)) <= 1e-15 * np.max(np.abs(want))
    # only the interior of `comp` changes
    mask = np.ones(a.shape, dtype=bool)
    mask[h:N + h, h:N + h, h:N + h, comp] = False
    assert np.array_equal(got[mask], a[mask])


def _stats(vel, mass):
    v2 = np.sum(vel * vel, axis=1)
    mv = mass[:, None] * vel
    return {"v2": v2, "mvx": mv[:, 0], "mvy": mv[:, 1], "mvz": mv[:, 2], "m": mass, "ekin": 0.5 * mass * v2}


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 1e-5)])
@pytest.mark.parametrize("per_particle_mass", [False, True])
def test_updateVel_outputs_match_oracle(prec, tol, per_particle_mass):
    """updateVel(updateVel_cicKick, dt, {phi}, output, reduce_type) on a 32^3 potential:
    sqrt(max |v|^2) and every slot under SUM, MIN, MAX and the _LOCAL forms"""
    N, n, dt = 32, 30000, 0.004
    announce(f"updateVel outputs {prec} pm={per_particle_mass}")
    rng = np.random.default_rng(17)
    pos = rng.random((n, 3))
    vel = 0.05 * rng.standard_normal((n, 3))
    mass = (rng.random(n) + 0.5) if per_particle_mass else None
    parts = lf.Particles(N, n, DT[prec], ids=True, per_particle_mass=per_particle_mass, mass=0.125)
    parts.addParticles(pos, vel, mass, ids=np.arange(n))
    d0 = parts.download()
    o0 = np.argsort(d0["id"])
    lat = lf.Lattice(3, N, 1)
    phi = lf.Field(lat, 1, DT[prec])
    g = torch.Generator(device="cpu").manual_seed(5)
    phi.t.copy_(torch.randn(phi.t.shape, generator=g, dtype=torch.float64).to(phi.t.dtype))
    phi_np = np.ascontiguousarray(real_np(phi))
    out = [0.0] * 6
    rtypes = [lf.SUM, lf.MAX, lf.SUM_LOCAL, lf.MIN, lf.SUM, lf.MAX_LOCAL]
    vmax = parts.updateVel(lf.updateVel_cicKick, dt, [phi], 1, None, out, rtypes, 6)
    # the oracle's kick of the same particles
    orc.ghost_fill(phi_np, N, 1)
    F = orc.force_field(phi_np, N, 1)
    v = np.ascontiguousarray(d0["vel"][o0].copy())
    orc.gather_kick(np.ascontiguousarray(d0["pos"][o0]), v, F, N, 1, dt_kick=dt)
    m = d0["mass"][o0] if per_particle_mass else np.full(n, 0.125)
    s = _stats(v, m)
    want = [s["v2"].sum(), s["mvx"].max(), s["mvy"].sum(), s["mvz"].min(), s["m"].sum(), s["ekin"].max()]
    scale = [np.abs(s["v2"]).sum(), np.abs(s["mvx"]).max(), np.abs(s["mvy"]).sum(), np.abs(s["mvz"]).max(),
             s["m"].sum(), s["ekin"].max()]
    for k in range(6):
        assert abs(out[k] - want[k]) <= tol * scale[k], (k, out[k], want[k])
    assert abs(vmax - np.sqrt(s["v2"].max())) <= tol * np.sqrt(s["v2"].max())
    # the kicked velocities the statistics were taken of
    d1 = parts.download()
    o1 = np.argsort(d1["id"])
    assert rel_l2(d1["vel"][o1] - d0["vel"][o0], v - d0["vel"][o0]) <= tol


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_updateVel_simple_and_move_outputs(prec):
    """updateVel_simple leaves the velocities as they are; moveParticles(move_particles_simple)
    drifts and reports the same statistics; a slot without a reduction type keeps its value;
    the starting values of MIN and MAX over no particles are the reference's +-9223372036854775807"""
    N, n = 16, 5000
    announce(f"updateVel_simple / moveParticles outputs {prec}")
    rng = np.random.default_rng(23)
    pos, vel = rng.random((n, 3)), rng.standard_normal((n, 3))
    parts = lf.Particles(N, n, DT[prec], ids=True, mass=2.0)
    parts.addParticles(pos, vel, ids=np.arange(n))
    d0 = parts.download()
    out = [0.0, 0.0, 7.5]
    vmax = parts.updateVel(lf.updateVel_simple, 0.1, None, 0, None, out, [lf.MIN, lf.SUM, 0], 3)
    s = _stats(d0["vel"], np.full(n, 2.0))
    tol = 1e-12
    assert abs(vmax - np.sqrt(s["v2"].max())) <= tol * vmax
    assert abs(out[0] - s["v2"].min()) <= tol * s["v2"].max() and out[2] == 7.5
    assert abs(out[1] - s["mvx"].sum()) <= tol * np.abs(s["mvx"]).sum()
    d1 = parts.download()
    assert np.array_equal(d1["vel"], d0["vel"])
    mo = [0.0]
    parts.moveParticles(lf.move_particles_simple, 0.5 / N, None, 0, None, mo, [lf.SUM], 1)
    assert abs(mo[0] - s["v2"].sum()) <= tol * s["v2"].sum()
    d2 = parts.download()
    o0, o2 = np.argsort(d0["id"]), np.argsort(d2["id"])
    moved = (d0["pos"][o0] + 0.5 / N * d0["vel"][o0]) % 1.0
    dpos = np.abs(d2["pos"][o2] - moved)
    dpos = np.minimum(dpos, 1.0 - dpos)
    # FP32: the offset is rounded to FP32 (2^-24 cell) and dt v is an FP32 product
    assert np.max(dpos) <= (1e-12 if prec == "f64" else 1e-7)
    with pytest.raises(lf.Lf2miError):
        parts.updateVel(lambda *a: 0.0, 0.1)
    empty = lf.Particles(N, 16, DT[prec])
    r = empty.reduce([lf.PQ_V2, lf.PQ_V2, lf.PQ_V2], [lf.MIN, lf.MAX, lf.SUM])
    assert r == [9223372036854775807.0, -9223372036854775807.0, 0.0]


def test_fixed_point_deposit_with_masses_of_order_one_over_np():
    """the scalar int64 fixed-point deposit with per-particle masses m ~ 1/Np (mean density
    1): the fixed point is scaled by the particles' bound (2^(62-e) with Np max m N^3 < 2^e),
    so it resolves masses far below one unit of 2^-32; rel-L2 <= 1e-6 against the oracle and
    bitwise reproducible after a re-bucketing"""
    N, n = 32, 2_000_000
    announce("fixed64 scalar deposit, masses ~ 1/Np")
    rng = np.random.default_rng(29)
    pos = rng.random((n, 3))
    mass = (1.0 + 0.5 * rng.random(n)) / n
    parts = lf.Particles(N, n, lf.F32, ids=True, per_particle_mass=True)
    parts.addParticles(pos, np.zeros((n, 3)), mass, ids=np.arange(n))
    d = parts.download()
    o = np.argsort(d["id"])
    lat = lf.Lattice(3, N, 1)
    rho = lf.Field(lat, 1, lf.F32)
    lf.projection_init(rho)
    parts.deposit(lf.PROJ_SCALAR, rho, lf.DEP_PERIODIC, lf.ACC_FIXED64)
    got = real_np(rho).copy()
    ref = orc.real_field(N, 1, 1)
    orc.scalar_project(np.ascontiguousarray(d["pos"][o]), ref, N, 1, mass=np.ascontiguousarray(d["mass"][o]))
    orc.scalar_comm(ref, N, 1)
    s = slice(1, N + 1)
    err = rel_l2(got[s, s, s], ref[s, s, s])
    print(f"fixed64, masses ~ 1/Np: rel-L2 {err:.2e}")
    assert err <= 1e-6
    # a zero drift re-buckets the particles into the other buffer: the same bits
    parts.gather_kick_drift(None, 0.0, 0.0)
    rho2 = lf.Field(lat, 1, lf.F32)
    lf.projection_init(rho2)
    parts.deposit(lf.PROJ_SCALAR, rho2, lf.DEP_PERIODIC, lf.ACC_FIXED64)
    assert torch.equal(rho.t, rho2.t)
