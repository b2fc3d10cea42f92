"""GPU parity: binning, CIC deposits (+ folds), fixed-point determinism, gather + kick + drift.

Oracle: oracle/ (FP64).  Both sides consume identical FP64 positions: the FP32
container represents a particle as (cell, FP32 offset); its download returns the
exact FP64 position (cell + offset)/N that the oracle is then given.  Bars:
  * cell keys and per-cell counts: bit-exact (integer work);
  * drift on identical velocities: cells bit-exact, positions within one FP32
    offset rounding (2^-24 cell);
  * FP64 plain path deposits / gather: <= 1e-12 relative;
  * FP32 deposits rel-L2 <= 1e-6, gathered accelerations rel-L2 <= 1e-5.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import latfield2_amd as lf  # noqa: E402
import oracle as orc  # noqa: E402
from gpu_util import announce, real_np, rel_l2  # noqa: E402

pytestmark = pytest.mark.gpu
DT = {"f32": lf.F32, "f64": lf.F64}


def _random_particles(n, seed, clustered=False):
    rng = np.random.default_rng(seed)
    pos = rng.random((n, 3))
    if clustered:
        pos[: n // 2] = (0.3 + 0.02 * rng.standard_normal((n // 2, 3))) % 1.0
    vel = rng.standard_normal((n, 3))
    mass = rng.random(n) + 0.5
    return pos, vel, mass


def _upload(N, pos, vel, prec, mass=None, cap=None):
    parts = lf.Particles(N, cap or len(pos), DT[prec], ids=True, per_particle_mass=mass is not None)
    kept = parts.addParticles(pos, vel, mass, ids=np.arange(len(pos)))
    return parts, kept


def _by_id(d):
    order = np.argsort(d["id"])
    return {k: v[order] for k, v in d.items()}


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_binning_bit_exact_with_edge_cases(prec):
    N = 32
    announce(f"binning {prec} N={N}")
    pos, vel, _ = _random_particles(20000, 1)
    pos[0] = [0.0, 0.0, 0.0]
    pos[1] = [1.0 - 2 ** -53, 0.5, 31 / 32]
    pos[2] = [1.0, 0.25, 0.0]       # pos == box: the reference's key is 0
    pos[3] = [-1e-9, 0.1, 0.1]      # negative coordinate: dropped by addParticle_global
    pos[4] = [5 / 32, 7 / 32 - 2 ** -60, 0.5]
    parts, kept = _upload(N, pos, vel, prec)
    assert kept == len(pos) - 1
    d = _by_id(parts.download())
    keys = orc.bin_keys(pos, N)
    valid = keys >= 0
    assert np.array_equal(d["id"], np.nonzero(valid)[0])
    assert np.array_equal(d["cell"], keys[valid])
    cnt = parts.cell_counts().cpu().numpy()
    assert np.array_equal(cnt, orc.cell_counts(pos, N))
    if prec == "f32":
        # the represented position is within one FP32 offset rounding of the input
        # (pos == 1.0 is represented as 0.0, its periodic image)
        p = pos[valid]
        diff = np.abs(d["pos"] - p)
        diff = np.minimum(diff, 1.0 - diff)
        assert np.max(diff) * N <= 2.0 ** -24
    else:
        assert np.array_equal(d["pos"], pos[valid])


def test_bin_positions_without_a_container_bit_exact():
    """lf2mi_bin_positions: getPartCoord keys and the per-cell histogram of raw FP64
    positions, the edge cases of the container test included"""
    N = 32
    announce(f"bin_positions N={N}")
    pos, _, _ = _random_particles(30000, 4)
    pos[0] = [0.0, 0.0, 0.0]
    pos[1] = [1.0 - 2 ** -53, 0.5, 31 / 32]
    pos[2] = [1.0, 0.25, 0.0]        # pos == box: key 0 along x
    pos[3] = [-1e-9, 0.1, 0.1]       # negative: key -1, not counted
    pos[4] = [5 / 32, 7 / 32 - 2 ** -60, 0.5]
    pos[5] = [1.25, 0.5, 0.5]        # beyond the box: floor(pos N) % N wraps
    keys, counts = lf.bin_positions(torch.as_tensor(pos, device="cuda"), N)
    torch.cuda.synchronize()
    assert np.array_equal(keys.cpu().numpy(), orc.bin_keys(pos, N))
    assert np.array_equal(counts.cpu().numpy(), orc.cell_counts(pos, N))
    k2, c2 = lf.bin_positions(torch.as_tensor(pos[:0], device="cuda"), N)  # empty: zero counts
    assert k2.numel() == 0 and int(c2.abs().sum()) == 0


def _oracle_deposits(pos, vel, mass, N, h, fold):
    rho = orc.real_field(N, h, 1)
    B = orc.real_field(N, h, 3)
    T = orc.real_field(N, h, 6)
    orc.scalar_project(pos, rho, N, h, mass=mass)
    orc.vector_project(pos, vel, B, N, h, mass=mass)
    orc.vecvec_project(pos, vel, T, N, h, mass=mass)
    if fold:
        orc.scalar_comm(rho, N, h)
        orc.vector_comm(B, N, h)
        orc.vecvec_comm(T, N, h)
    return [rho, B, T]


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 1e-6)])
@pytest.mark.parametrize("N,h,n,clustered,masses", [(16, 1, 5000, False, False), (32, 2, 40000, True, True),
                                                     (24, 1, 3000, True, False),
                                                     (64, 1, 6000, False, False),     # ~12 per brick: direct path
                                                     (16, 1, 120000, True, True),     # ~15k per brick: many items
                                                     # brick edges other than 8 (N % 8 != 0): B = 5, 6, 7
                                                     (20, 1, 8000, True, True),
                                                     (12, 2, 20000, True, False),
                                                     (14, 1, 500, False, False)])
@pytest.mark.parametrize("mode", ["unfolded", "unfolded+comm", "periodic"])
def test_deposits_match_oracle(prec, tol, N, h, n, clustered, masses, mode):
    announce(f"deposits {prec} N={N} h={h} n={n} clustered={clustered} masses={masses} mode={mode}")
    pos, vel, mass = _random_particles(n, N + n, clustered)
    parts, _ = _upload(N, pos, vel, prec, mass if masses else None)
    d = parts.download()  # the positions/velocities the container represents
    lat = lf.Lattice(3, N, h)
    fields = [lf.Field(lat, C, DT[prec]) for C in (1, 3, 6)]
    kinds = [lf.PROJ_SCALAR, lf.PROJ_VECTOR, lf.PROJ_VECVEC]
    comms = [lf.scalarProjectionCIC_comm, lf.vectorProjectionCIC_comm, lf.VecVecProjectionCIC_comm]
    for f, kd, cm in zip(fields, kinds, comms):
        lf.projection_init(f)
        parts.deposit(kd, f, lf.DEP_PERIODIC if mode == "periodic" else lf.DEP_UNFOLDED)
        if mode == "unfolded+comm":
            cm(f)
    torch.cuda.synchronize()
    refs = _oracle_deposits(d["pos"], d["vel"], d["mass"] if masses else None, N, h, fold=(mode != "unfolded"))
    G = N + 2 * h
    for f, ref in zip(fields, refs):
        got = real_np(f)
        if mode == "periodic":
            got = got[h:G - h, h:G - h, h:G - h]
            ref = ref[h:G - h, h:G - h, h:G - h]
        assert rel_l2(got, ref) <= tol, (mode, f.C, rel_l2(got, ref))


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 1e-6)])
@pytest.mark.parametrize("N,h,n,clustered,masses", [(16, 1, 5000, False, False), (32, 2, 40000, True, True),
                                                     (64, 1, 6000, False, False),
                                                     (16, 1, 120000, True, True),
                                                     (20, 1, 8000, True, True),
                                                     (14, 1, 500, False, False)])
@pytest.mark.parametrize("mode", ["unfolded", "unfolded+comm", "periodic"])
def test_cic_ngp_deposits_match_oracle(prec, tol, N, h, n, clustered, masses, mode):
    """vectorProjectionCICNGP_* and symtensorProjectionCICNGP_* (projections.hpp:405-976)"""
    announce(f"CIC-NGP deposits {prec} N={N} h={h} n={n} clustered={clustered} masses={masses} mode={mode}")
    pos, vel, mass = _random_particles(n, 7 * N + n, clustered)
    parts, _ = _upload(N, pos, vel, prec, mass if masses else None)
    d = parts.download()
    lat = lf.Lattice(3, N, h)
    B, T = lf.Field(lat, 3, DT[prec]), lf.Field(lat, 6, DT[prec])
    target = lf.DEP_PERIODIC if mode == "periodic" else lf.DEP_UNFOLDED
    for f, kd, cm in ((B, lf.PROJ_VECTOR_NGP, lf.vectorProjectionCICNGP_comm),
                      (T, lf.PROJ_SYMTENSOR_NGP, lf.symtensorProjectionCICNGP_comm)):
        lf.projection_init(f)
        parts.deposit(kd, f, target)
        if mode == "unfolded+comm":
            cm(f)
    torch.cuda.synchronize()
    m = d["mass"] if masses else None
    rB, rT = orc.real_field(N, h, 3), orc.real_field(N, h, 6)
    orc.vector_ngp_project(d["pos"], d["vel"], rB, N, h, mass=m)
    orc.symtensor_ngp_project(d["pos"], d["vel"], rT, N, h, mass=m)
    if mode != "unfolded":
        orc.vector_ngp_comm(rB, N, h)
        orc.symtensor_ngp_comm(rT, N, h)
    G = N + 2 * h
    for f, ref in ((B, rB), (T, rT)):
        got = real_np(f)
        if mode == "periodic":
            got = got[h:G - h, h:G - h, h:G - h]
            ref = ref[h:G - h, h:G - h, h:G - h]
        for c in range(f.C):
            assert rel_l2(got[..., c], ref[..., c]) <= tol, (mode, f.C, c, rel_l2(got[..., c], ref[..., c]))


@pytest.mark.parametrize("N,n,clustered,masses", [
    (64, 360000, False, False),   # ~700 per brick: most of them read again in every component pass
    (64, 200000, True, True),     # a blob of dense bricks (item kernel) among sparse ones
    (128, 400000, True, False),   # several tiles per row of bricks: seams between tiles and across the edge
    (128, 400000, True, True),
    (96, 300000, True, True)])    # N not a power of two (12 bricks per row: 3 tiles of the multi-component kernel)
def test_periodic_tile_deposits_match_oracle(N, n, clustered, masses):
    """the periodic FP32 deposits on lattices whose rows of bricks split into tiles of 8:
    every kind (and the scalar kind with per-particle masses) against the FP64 oracle"""
    announce(f"tile deposits N={N} n={n} clustered={clustered} masses={masses}")
    pos, vel, mass = _random_particles(n, 11 * N + n, clustered)
    parts, _ = _upload(N, pos, vel, "f32", mass if masses else None)
    d = parts.download()
    m = d["mass"] if masses else None
    h = 1
    lat = lf.Lattice(3, N, h)
    cases = ((lf.PROJ_SCALAR, 1, lambda r: orc.scalar_project(d["pos"], r, N, h, mass=m), orc.scalar_comm),
             (lf.PROJ_VECTOR, 3, lambda r: orc.vector_project(d["pos"], d["vel"], r, N, h, mass=m), orc.vector_comm),
             (lf.PROJ_VECVEC, 6, lambda r: orc.vecvec_project(d["pos"], d["vel"], r, N, h, mass=m), orc.vecvec_comm),
             (lf.PROJ_VECTOR_NGP, 3, lambda r: orc.vector_ngp_project(d["pos"], d["vel"], r, N, h, mass=m),
              orc.vector_ngp_comm),
             (lf.PROJ_SYMTENSOR_NGP, 6, lambda r: orc.symtensor_ngp_project(d["pos"], d["vel"], r, N, h, mass=m),
              orc.symtensor_ngp_comm))
    s = slice(h, N + h)
    for kind, C, project, comm in cases:
        f = lf.Field(lat, C, lf.F32)
        lf.projection_init(f)
        parts.deposit(kind, f, lf.DEP_PERIODIC)
        torch.cuda.synchronize()
        ref = orc.real_field(N, h, C)
        project(ref)
        comm(ref, N, h)
        G = N + 2 * h
        got = real_np(f).reshape(G, G, G, C)
        for c in range(C):
            err = rel_l2(got[s, s, s, c], ref[s, s, s, c])
            assert err <= 1e-6, (kind, c, err)
        # the ghost layer is not written by a periodic deposit
        ghosts = got.copy()
        ghosts[s, s, s] = 0.0
        assert not ghosts.any(), kind


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("N,n,masses", [(64, 200000, True), (128, 400000, False), (32, 40000, False)])
def test_periodic_set_deposit_replaces_the_field(prec, N, n, masses):
    """LF2MI_DEP_PERIODIC_SET == projection_init + periodic deposit, whatever the field held:
    the interior equals the deposit into a cleared field (bitwise with the fixed point, FP32
    atomics to rounding), the ghosts are zero"""
    announce(f"periodic set deposit {prec} N={N} n={n} masses={masses}")
    pos, vel, mass = _random_particles(n, 5 * N + n, clustered=True)
    parts, _ = _upload(N, pos, vel, prec, mass if masses else None)
    h = 1
    lat = lf.Lattice(3, N, h)
    accs = [lf.ACC_NATIVE] + ([lf.ACC_FIXED64] if prec == "f32" else [])
    for kind, C in ((lf.PROJ_SCALAR, 1), (lf.PROJ_VECTOR, 3), (lf.PROJ_VECVEC, 6), (lf.PROJ_VECTOR_NGP, 3),
                    (lf.PROJ_SYMTENSOR_NGP, 6)):
        for acc in accs:
            a, b = lf.Field(lat, C, DT[prec]), lf.Field(lat, C, DT[prec])
            a.t.fill_(0.25)
            parts.deposit(kind, a, lf.DEP_PERIODIC_SET, acc)
            lf.projection_init(b)
            parts.deposit(kind, b, lf.DEP_PERIODIC, acc)
            torch.cuda.synchronize()
            G = N + 2 * h
            ga, gb = real_np(a).reshape(G, G, G, C), real_np(b).reshape(G, G, G, C)
            s = slice(h, N + h)
            ghosts = ga.copy()
            ghosts[s, s, s] = 0.0
            assert not ghosts.any(), (kind, acc)
            if acc == lf.ACC_FIXED64:
                assert np.array_equal(ga, gb), kind
            else:
                err = rel_l2(ga, gb)
                assert err <= (1e-12 if prec == "f64" else 1e-6), (kind, acc, err)
    empty = lf.Particles(N, 10, DT[prec])
    f = lf.Field(lat, 3, DT[prec])
    f.t.fill_(0.25)
    empty.deposit(lf.PROJ_VECTOR, f, lf.DEP_PERIODIC_SET)
    torch.cuda.synchronize()
    assert not bool(f.t.any())


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("N", [16, 64])
def test_deposits_of_an_empty_container_add_nothing(prec, N):
    announce(f"empty container deposits {prec} N={N}")
    parts = lf.Particles(N, 10, DT[prec])
    lat = lf.Lattice(3, N, 1)
    for kind, C in ((lf.PROJ_SCALAR, 1), (lf.PROJ_VECTOR, 3), (lf.PROJ_VECVEC, 6), (lf.PROJ_VECTOR_NGP, 3),
                    (lf.PROJ_SYMTENSOR_NGP, 6)):
        f = lf.Field(lat, C, DT[prec])
        f.t.fill_(0.25)
        for target in (lf.DEP_UNFOLDED, lf.DEP_PERIODIC):
            parts.deposit(kind, f, target)
        torch.cuda.synchronize()
        assert bool((f.t == 0.25).all()), kind
    with pytest.raises(lf.Lf2miError):
        parts.deposit(5, lf.Field(lat, 6, DT[prec]), lf.DEP_PERIODIC)  # no such kind


def test_fixed_point_deposit_is_reproducible_and_accurate():
    N = 32
    announce(f"fixed-point deposit N={N}")
    pos, vel, _ = _random_particles(60000, 3, clustered=True)
    parts, _ = _upload(N, pos, vel, "f32")
    lat = lf.Lattice(3, N, 1)
    a, b, c = (lf.Field(lat, 1, lf.F32) for _ in range(3))
    for f in (a, b):
        lf.projection_init(f)
        parts.deposit(lf.PROJ_SCALAR, f, lf.DEP_PERIODIC, lf.ACC_FIXED64)
    lf.projection_init(c)
    parts.deposit(lf.PROJ_SCALAR, c, lf.DEP_PERIODIC, lf.ACC_NATIVE)
    torch.cuda.synchronize()
    assert torch.equal(a.t, b.t)
    d = parts.download()
    ref = orc.real_field(N, 1, 1)
    orc.scalar_project(d["pos"], ref, N, 1)
    orc.scalar_comm(ref, N, 1)
    assert rel_l2(real_np(a)[1:N + 1, 1:N + 1, 1:N + 1], ref[1:N + 1, 1:N + 1, 1:N + 1]) <= 1e-6
    assert rel_l2(real_np(c)[1:N + 1, 1:N + 1, 1:N + 1], ref[1:N + 1, 1:N + 1, 1:N + 1]) <= 1e-6


_FX_KINDS = [(lf.PROJ_SCALAR, 1, "scalar_project", "scalar_comm"), (lf.PROJ_VECTOR, 3, "vector_project", "vector_comm"),
             (lf.PROJ_VECVEC, 6, "vecvec_project", "vecvec_comm"),
             (lf.PROJ_VECTOR_NGP, 3, "vector_ngp_project", "vector_ngp_comm"),
             (lf.PROJ_SYMTENSOR_NGP, 6, "symtensor_ngp_project", "symtensor_ngp_comm")]


@pytest.mark.parametrize("kind,C,proj,comm", _FX_KINDS)
@pytest.mark.parametrize("masses", [False, True])
def test_fixed_point_deposits_are_order_independent(kind, C, proj, comm, masses):
    """LF2MI_ACC_FIXED64 for every kind: bitwise equal after the particles were re-bucketed
    into another order (a zero drift), and within 1e-6 of the FP64 oracle"""
    N = 24
    announce(f"fixed-point deposit kind={kind} masses={masses}")
    pos, vel, mass = _random_particles(50000, 13 + kind, clustered=True)
    parts, _ = _upload(N, pos, vel, "f32", mass if masses else None)
    lat = lf.Lattice(3, N, 1)
    a, b = lf.Field(lat, C, lf.F32), lf.Field(lat, C, lf.F32)
    lf.projection_init(a)
    parts.deposit(kind, a, lf.DEP_PERIODIC, lf.ACC_FIXED64)
    d = parts.download()
    parts.gather_kick_drift(None, dt_kick=0.0, dt_drift=0.0)  # re-bucket: another order in the buckets
    lf.projection_init(b)
    parts.deposit(kind, b, lf.DEP_PERIODIC, lf.ACC_FIXED64)
    torch.cuda.synchronize()
    assert torch.equal(a.t, b.t)
    ref = orc.real_field(N, 1, C)
    m = d["mass"] if masses else None
    if kind == lf.PROJ_SCALAR:
        getattr(orc, proj)(d["pos"], ref, N, 1, mass=m)
    else:
        getattr(orc, proj)(d["pos"], d["vel"], ref, N, 1, mass=m)
    getattr(orc, comm)(ref, N, 1)
    s_ = slice(1, N + 1)
    got = real_np(a)[s_, s_, s_]
    for c in range(C):
        assert rel_l2(got[..., c], ref[s_, s_, s_, c]) <= 1e-6, c


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_drift_cells_bit_exact_on_identical_velocities(prec):
    """dt_kick = 0: both sides drift the same FP64 position with the same velocity"""
    N = 64
    announce(f"drift bit-exact {prec} N={N}")
    rng = np.random.default_rng(17)
    pos = rng.random((50000, 3))
    pos[:100] = np.floor(pos[:100] * N) / N  # on cell faces
    vel = 0.05 * rng.standard_normal((50000, 3))
    vel[:50] = 0.0
    parts, _ = _upload(N, pos, vel, prec)
    d0 = _by_id(parts.download())
    phi = lf.Field(lf.Lattice(3, N, 1), 1, DT[prec])
    parts.gather_kick_drift(phi, dt_kick=0.0, dt_drift=0.01)
    torch.cuda.synchronize()
    d1 = _by_id(parts.download())
    p = np.ascontiguousarray(d0["pos"])
    orc.drift(p, d0["vel"], 0.01)
    keys = orc.bin_keys(p, N)
    assert np.array_equal(d1["cell"], keys)
    diff = np.abs(d1["pos"] - p)
    diff = np.minimum(diff, 1.0 - diff)
    bound = 0.0 if prec == "f64" else 2.0 ** -24 / N
    assert np.max(diff) <= bound
    assert np.array_equal(parts.cell_counts().cpu().numpy(), orc.cell_counts(p, N))


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 1e-5)])
@pytest.mark.parametrize("N,n", [(32, 30000), (64, 4000), (16, 100000),
                                 (20, 8000), (12, 1500), (14, 20000)])   # B = 5, 6, 7
def test_gather_kick_matches_oracle(prec, tol, N, n):
    """item path (dense bricks, several items per brick) and chunk path (runs of sparse
    bricks), for brick edges 8 and 5-7"""
    announce(f"gather+kick {prec} N={N} n={n}")
    pos, vel, _ = _random_particles(n, 21 + N, clustered=True)
    vel *= 0.01
    parts, _ = _upload(N, pos, vel, prec)
    d0 = _by_id(parts.download())
    lat = lf.Lattice(3, N, 1)
    phi = lf.Field(lat, 1, DT[prec])
    g = torch.Generator(device="cpu").manual_seed(9)
    phi.t.copy_(torch.randn(phi.t.shape, generator=g, dtype=torch.float64).to(phi.t.dtype))
    phi_np = np.ascontiguousarray(real_np(phi))
    dt = 0.003
    vmax2 = parts.gather_kick_drift(phi, dt_kick=dt, dt_drift=0.0, want_vmax=True)
    torch.cuda.synchronize()
    d1 = _by_id(parts.download())
    orc.ghost_fill(phi_np, N, 1)  # the kernel wraps in the interior; ghosts are its periodic image
    F = orc.force_field(phi_np, N, 1)
    v = np.ascontiguousarray(d0["vel"].copy())
    acc, vm = orc.gather_kick(d0["pos"], v, F, N, 1, dt_kick=dt)
    assert rel_l2(d1["vel"] - d0["vel"], v - d0["vel"]) <= tol
    assert abs(vmax2 - vm) <= 1e-6 * vm
    # dt_drift = 0: positions and cells unchanged
    assert np.array_equal(d1["cell"], d0["cell"])
