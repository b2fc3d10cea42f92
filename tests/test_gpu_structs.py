"""GPU parity of user particle structs (Particles<part, ...> with any part type,
particles/LATfield2_Particles.hpp:131-344): the container keeps each particle's struct and
moves it through the re-bucketing; the deposits take the mass from part_info or from a
double at any byte offset of the struct (projections.hpp:108-126, :144-152), and the vector
and tensor deposits their vector from any Real[3] of it (oset[1]).

Bars: kept struct members bit-exact by ID after moves; positions and velocities as in
lf2mi_parts_download; deposits rel-L2 <= 1e-6 (FP32) / 1e-12 (FP64) against the FP64 oracle
with the masses each source selects.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import latfield2_amd as lf  # noqa: E402
import oracle as orc  # noqa: E402
from gpu_util import announce, real_np, rel_l2  # noqa: E402

pytestmark = pytest.mark.gpu
DT = {"f32": lf.F32, "f64": lf.F64}
TOL = {"f32": 1e-6, "f64": 1e-12}


def _dtype(real):
    return np.dtype([("ID", "<i8"), ("pos", real, 3), ("vel", real, 3), ("mass", "<f8"), ("w", "<f8"),
                     ("tag", "<i8")], align=True)


def _structs(n, seed, real):
    rng = np.random.default_rng(seed)
    a = np.zeros(n, dtype=_dtype(real))
    a["ID"] = np.arange(n) * 3 + 11
    pos = rng.random((n, 3))
    pos[: n // 2] = (0.4 + 0.03 * rng.standard_normal((n // 2, 3))) % 1.0
    a["pos"] = pos
    a["vel"] = rng.standard_normal((n, 3))
    a["mass"] = rng.random(n) + 0.5
    a["w"] = rng.random(n) * 2.0 + 0.1
    a["tag"] = rng.integers(0, 2 ** 62, n)
    return a


def _by_id(a):
    return a[np.argsort(a["ID"])]


def _deposit(parts, N, prec):
    lat = lf.Lattice(3, N, 1)
    f = lf.Field(lat, 1, DT[prec])
    f.zero()
    parts.deposit(lf.PROJ_SCALAR, f, lf.DEP_PERIODIC)
    torch.cuda.synchronize()
    s = slice(1, N + 1)
    return real_np(f)[s, s, s]


def _oracle(pos, mass, N):
    ref = orc.real_field(N, 1, 1)
    orc.scalar_project(np.ascontiguousarray(pos), ref, N, 1, mass=np.ascontiguousarray(mass))
    orc.scalar_comm(ref, N, 1)
    s = slice(1, N + 1)
    return ref[s, s, s]


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("real", ["<f8", "<f4"])
def test_structs_round_trip_and_mass_sources(prec, real):
    N, n = 24, 6000
    announce(f"structs {prec} Real={real} N={N} n={n}")
    a = _structs(n, 5, real)
    parts = lf.Particles(N, n, DT[prec], ids=True, mass=0.7)
    parts.set_payload(a.dtype.itemsize)
    assert parts.upload_structs(a) == n
    b = _by_id(parts.download_structs(a.dtype))
    ref = _by_id(a)
    for k in ("ID", "mass", "w", "tag"):
        assert np.array_equal(b[k], ref[k]), k
    d = parts.download()
    order = np.argsort(d["id"])
    assert np.array_equal(b["pos"], d["pos"][order].astype(a.dtype["pos"].base))
    # represented positions: the struct's within one FP32 offset rounding (exact in F64)
    diff = np.abs(d["pos"][order] - ref["pos"].astype(np.float64))
    diff = np.minimum(diff, 1.0 - diff)
    assert np.max(diff) * N <= (1.0001 * 2.0 ** -24 if prec == "f32" else 0.0)
    pos = np.ascontiguousarray(d["pos"][order])
    tol = TOL[prec]
    # default: no per-particle masses were asked for, so part_info's mass
    got = _deposit(parts, N, prec)
    assert rel_l2(got, _oracle(pos, np.full(n, 0.7), N)) <= tol
    # the masses uploaded with the structs (offset -1: the per-particle masses as they stand)
    parts.mass_from(lf.MASS_PART, -1)
    got = _deposit(parts, N, prec)
    assert rel_l2(got, _oracle(pos, ref["mass"], N)) <= tol
    # FROM_PART at the weight's offset, then at the mass member's
    parts.mass_from(lf.MASS_PART, a.dtype.fields["w"][1])
    got = _deposit(parts, N, prec)
    assert rel_l2(got, _oracle(pos, ref["w"], N)) <= tol
    parts.mass_from(lf.MASS_PART, a.dtype.fields["mass"][1])
    got = _deposit(parts, N, prec)
    assert rel_l2(got, _oracle(pos, ref["mass"], N)) <= tol
    # back to FROM_INFO
    parts.mass_from(lf.MASS_INFO)
    got = _deposit(parts, N, prec)
    assert rel_l2(got, _oracle(pos, np.full(n, 0.7), N)) <= tol


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_structs_follow_their_particles_through_moves(prec):
    N, n = 32, 20000
    announce(f"structs moves {prec} N={N} n={n}")
    a = _structs(n, 9, "<f8")
    a["vel"] *= 1.5 / (N * 0.01)  # ~1.5 cells per step
    parts = lf.Particles(N, n, DT[prec], ids=True)
    parts.set_payload(a.dtype.itemsize)
    parts.upload_structs(a)
    parts.mass_from(lf.MASS_PART, a.dtype.fields["w"][1])
    ref = _by_id(a)
    for step in range(4):
        parts.gather_kick_drift(None, dt_kick=0.0, dt_drift=0.01)
        torch.cuda.synchronize()
        b = _by_id(parts.download_structs(a.dtype))
        for k in ("ID", "mass", "w", "tag"):
            assert np.array_equal(b[k], ref[k]), (step, k)
        d = parts.download()
        order = np.argsort(d["id"])
        assert np.array_equal(b["pos"], d["pos"][order])
        assert np.array_equal(b["vel"], d["vel"][order])
        # the weights still drive the deposit after the moves
        got = _deposit(parts, N, prec)
        assert rel_l2(got, _oracle(np.ascontiguousarray(d["pos"][order]), ref["w"], N)) <= TOL[prec], step


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_float_mass_member_read_with_its_own_type(prec):
    """a `Real mass` member under SINGLE is a float: read as a float (the reference reads every
    per-particle mass as a double, projections.hpp:150)"""
    N, n = 16, 4000
    announce(f"structs float mass {prec}")
    dt = np.dtype([("ID", "<i8"), ("pos", "<f4", 3), ("vel", "<f4", 3), ("mass", "<f4"), ("pad", "<f4")], align=True)
    rng = np.random.default_rng(21)
    a = np.zeros(n, dtype=dt)
    a["ID"] = np.arange(n)
    a["pos"] = rng.random((n, 3))
    a["mass"] = rng.random(n) + 0.25
    parts = lf.Particles(N, n, DT[prec], ids=True, per_particle_mass=True, mass=0.7)
    parts.set_payload(dt.itemsize)
    assert parts.upload_structs(a) == n
    d = parts.download()
    order = np.argsort(d["id"])
    pos = np.ascontiguousarray(d["pos"][order])
    got = _deposit(parts, N, prec)
    assert rel_l2(got, _oracle(pos, a["mass"].astype(np.float64), N)) <= TOL[prec]
    # an explicit FROM_PART of the float member, then back to the global mass
    parts.mass_from(lf.MASS_INFO)
    parts.mass_from(lf.MASS_PART_F32, dt.fields["mass"][1])
    got = _deposit(parts, N, prec)
    assert rel_l2(got, _oracle(pos, a["mass"].astype(np.float64), N)) <= TOL[prec]
    parts.mass_from(lf.MASS_INFO)
    got = _deposit(parts, N, prec)
    assert rel_l2(got, _oracle(pos, np.full(n, 0.7), N)) <= TOL[prec]
    with pytest.raises(lf.Lf2miError):
        parts.mass_from(lf.MASS_PART_F32, dt.itemsize - 2)  # past the struct


_VEC_KINDS = [(lf.PROJ_VECTOR, 3, "vector_project"), (lf.PROJ_VECVEC, 6, "vecvec_project"),
              (lf.PROJ_VECTOR_NGP, 3, "vector_ngp_project"), (lf.PROJ_SYMTENSOR_NGP, 6, "symtensor_ngp_project")]


def _vector_deposits(parts, N, prec, pos, vec, tol, what):
    lat = lf.Lattice(3, N, 1)
    s = slice(1, N + 1)
    for kind, C, fn in _VEC_KINDS:
        f = lf.Field(lat, C, DT[prec])
        f.zero()
        parts.deposit(kind, f, lf.DEP_PERIODIC)
        torch.cuda.synchronize()
        ref = orc.real_field(N, 1, C)
        getattr(orc, fn)(np.ascontiguousarray(pos), np.ascontiguousarray(vec), ref, N, 1, mass_global=0.7)
        getattr(orc, fn.replace("project", "comm"))(ref, N, 1)
        got = real_np(f)[s, s, s]
        for c in range(C):
            assert rel_l2(got[..., c], ref[s, s, s, c]) <= tol, (what, fn, c)


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("real", ["<f8", "<f4"])
def test_vector_projections_of_any_struct_member(prec, real):
    """a projection's oset[1] (projections.hpp:453-459, :741-747, :1019-1025, :1394-1400)"""
    N, n = 20, 5000
    announce(f"structs vector member {prec} Real={real} N={N} n={n}")
    dt = np.dtype([("ID", "<i8"), ("pos", real, 3), ("vel", real, 3), ("mom", real, 3), ("mass", "<f8")], align=True)
    rng = np.random.default_rng(17)
    a = np.zeros(n, dtype=dt)
    a["ID"] = np.arange(n)
    a["pos"] = rng.random((n, 3))
    a["vel"] = rng.standard_normal((n, 3))
    a["mom"] = rng.standard_normal((n, 3)) * 3.0 + 1.0
    parts = lf.Particles(N, n, DT[prec], ids=True, mass=0.7)
    parts.set_payload(dt.itemsize)
    assert parts.upload_structs(a) == n
    parts.mass_from(lf.MASS_INFO)
    code = lf.F32 if real == "<f4" else lf.F64
    # a drift re-buckets the particles (and their structs) first
    parts.gather_kick_drift(None, dt_kick=0.0, dt_drift=0.3 / N)
    parts.vector_from(dt.fields["mom"][1], code)
    d = parts.download()
    order = np.argsort(d["id"])
    pos = d["pos"][order]
    mom = _by_id(parts.download_structs(dt))["mom"].astype(np.float64)
    _vector_deposits(parts, N, prec, pos, mom, TOL[prec], "mom")
    # a kick changes the container's velocity; the vel offset selects that velocity, not the
    # uploaded copy in the struct
    lat = lf.Lattice(3, N, 1)
    phi = lf.Field(lat, 1, DT[prec])
    g = torch.Generator().manual_seed(3)
    phi.t.copy_(torch.randn(phi.t.shape, generator=g, dtype=torch.float64).to(phi.t.dtype))
    parts.gather_kick_drift(phi, dt_kick=0.05, dt_drift=0.0)
    parts.vector_from(dt.fields["vel"][1], code)
    d = parts.download()
    order = np.argsort(d["id"])
    assert not np.allclose(d["vel"][order], a["vel"].astype(np.float64))
    _vector_deposits(parts, N, prec, d["pos"][order], d["vel"][order], TOL[prec], "vel")
    parts.vector_from(-1)
    _vector_deposits(parts, N, prec, d["pos"][order], d["vel"][order], TOL[prec], "default")


def test_struct_errors():
    announce("structs errors")
    a = _structs(100, 1, "<f8")
    parts = lf.Particles(16, 100, lf.F32, ids=True)
    with pytest.raises(lf.Lf2miError):
        parts.set_payload(12)  # not a multiple of 8
    parts.set_payload(a.dtype.itemsize)
    with pytest.raises(lf.Lf2miError):
        parts.mass_from(lf.MASS_PART, a.dtype.itemsize - 4)  # past the struct
    b = np.zeros(100, dtype=[("ID", "<i8"), ("pos", "<f8", 3), ("vel", "<f8", 3)])
    with pytest.raises(lf.Lf2miError):
        parts.upload_structs(b)  # stride differs from the payload
    parts.upload_structs(a)
    with pytest.raises(lf.Lf2miError):
        parts.set_payload(a.dtype.itemsize)  # only before the first upload
    with pytest.raises(lf.Lf2miError):
        parts.vector_from(a.dtype.itemsize - 8, lf.F64)  # 24 bytes past the struct
    with pytest.raises(lf.Lf2miError):
        parts.vector_from(0, 7)  # bad Real type
    plain = lf.Particles(16, 100, lf.F32, ids=True)
    with pytest.raises(lf.Lf2miError):
        plain.mass_from(lf.MASS_PART, 0)  # no masses and no payload
    with pytest.raises(lf.Lf2miError):
        plain.vector_from(0, lf.F32)  # no payload
