"""GPU parity: ghost fill, CIC folds, R2C/C2R and C2C FFT, Green's function, fused Poisson solve.

Oracle: oracle/ (FP64 restatement, pinned in tests/test_oracle.py).  Bars:
  * ghost fill and folds are copies / single adds: bit-exact in both precisions;
  * FP64 FFT/Green/Poisson: |err| <= 1e-12 relative (max over the field);
  * FP32: rel-L2 <= 2e-6 against the FP64 oracle;
  * the reference KATs (wave_test*.cpp, tolerance 1e-10) hold on the FP64 path.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import latfield2_amd as lf  # noqa: E402
import oracle as orc  # noqa: E402
from gpu_util import announce, kfield_np, real_np, rel_l2  # noqa: E402

pytestmark = pytest.mark.gpu

DT = {"f32": lf.F32, "f64": lf.F64}


def _field(N, h, C, prec, vals=None, seed=0):
    f = lf.Field(lf.Lattice(3, N, h), C, DT[prec])
    if vals is None:
        g = torch.Generator(device="cpu").manual_seed(seed)
        vals = torch.randn(f.t.shape, generator=g, dtype=torch.float64)
    f.t.copy_(torch.as_tensor(vals).to(f.t.dtype))
    return f


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("N,h,C", [(16, 1, 1), (12, 2, 3), (8, 3, 6)])
def test_ghost_fill_bit_exact(prec, N, h, C):
    announce(f"ghost_fill {prec} N={N} h={h} C={C}")
    f = _field(N, h, C, prec, seed=N + h + C)
    inner = real_np(f)[h:N + h, h:N + h, h:N + h].copy()
    f.updateHalo()
    torch.cuda.synchronize()
    ref = np.pad(inner, ((h, h), (h, h), (h, h), (0, 0)), mode="wrap")
    assert np.array_equal(real_np(f), ref)


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("kind,C", [(lf.PROJ_SCALAR, 1), (lf.PROJ_VECTOR, 3), (lf.PROJ_VECVEC, 6)])
@pytest.mark.parametrize("N,h", [(8, 1), (10, 2)])
def test_projection_comm_bit_exact(prec, kind, C, N, h):
    announce(f"projection_comm {prec} kind={kind} N={N} h={h}")
    # integer values: every partial sum is exact in FP32 and FP64, so the fold order of
    # the reference is the only thing that can make the results differ
    rng = np.random.default_rng(kind * 10 + N)
    G = N + 2 * h
    f = _field(N, h, C, prec, vals=rng.integers(-1000, 1000, size=(G, G, G, C)).astype(np.float64))
    a = real_np(f).copy()
    comm = {lf.PROJ_SCALAR: lf.scalarProjectionCIC_comm, lf.PROJ_VECTOR: lf.vectorProjectionCIC_comm,
            lf.PROJ_VECVEC: lf.VecVecProjectionCIC_comm}[kind]
    comm(f)
    torch.cuda.synchronize()
    ref = a.copy()
    {lf.PROJ_SCALAR: orc.scalar_comm, lf.PROJ_VECTOR: orc.vector_comm, lf.PROJ_VECVEC: orc.vecvec_comm}[kind](ref, N, h)
    got = real_np(f)
    assert np.array_equal(got, ref)
    assert not np.array_equal(ref, a)  # the fold changed the field


def _fft_case(N, h, kh, C, prec, seed):
    lat = lf.Lattice(3, N, h)
    r = _field(N, h, C, prec, seed=seed)
    k = lf.KField(lat, kh, C, DT[prec])
    plan = lf.PlanFFT(r, k)
    return r, k, plan


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 2e-6)])
@pytest.mark.parametrize("N,h,kh,C", [(8, 1, 0, 1), (16, 1, 0, 1), (32, 2, 1, 3), (64, 1, 0, 1), (128, 1, 0, 1),
                                      (16, 1, 2, 6),
                                      # three or five times a power of two (radix-3/5 last pass)
                                      (12, 1, 0, 1), (24, 1, 1, 2), (96, 1, 0, 1), (192, 2, 1, 1),
                                      (10, 1, 0, 1), (40, 1, 1, 2), (80, 1, 0, 1), (160, 1, 0, 1)])
def test_fft_forward_matches_oracle(prec, tol, N, h, kh, C):
    announce(f"fft forward {prec} N={N} h={h} kh={kh} C={C}")
    r, k, plan = _fft_case(N, h, kh, C, prec, seed=N + C)
    plan.execute(lf.FFT_FORWARD)
    torch.cuda.synchronize()
    rin = np.ascontiguousarray(real_np(r))
    ref = orc.fft_r2c(rin, N, h, kh)
    got = kfield_np(k)
    assert rel_l2(got, ref) <= tol if prec == "f32" else np.max(np.abs(got - ref)) <= tol * np.max(np.abs(ref))
    # input preserved, k ghosts untouched
    assert np.array_equal(real_np(r), rin)
    if kh:
        assert np.all(got[:kh] == 0) and np.all(got[:, :, :kh] == 0)


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 2e-6)])
@pytest.mark.parametrize("N,h,kh,C", [(8, 1, 0, 1), (32, 1, 1, 1), (64, 2, 0, 3), (256, 1, 0, 1),
                                      (12, 1, 0, 1), (48, 1, 1, 3), (96, 1, 0, 1), (20, 1, 0, 1), (160, 1, 1, 1)])
def test_fft_backward_matches_oracle(prec, tol, N, h, kh, C):
    announce(f"fft backward {prec} N={N} h={h} kh={kh} C={C}")
    lat = lf.Lattice(3, N, h)
    k = lf.KField(lat, kh, C, DT[prec])
    g = torch.Generator(device="cpu").manual_seed(N * 7 + C)
    k.t.copy_(torch.randn(k.t.shape, generator=g, dtype=torch.float64).to(k.t.dtype))
    r = lf.Field(lat, C, DT[prec])
    plan = lf.PlanFFT(r, k)
    kin = kfield_np(k).copy()
    plan.execute(lf.FFT_BACKWARD)
    torch.cuda.synchronize()
    ref = orc.real_field(N, h, C)
    orc.fft_c2r(np.ascontiguousarray(kin), ref, N, h, kh)
    got = real_np(r)
    if prec == "f32":
        assert rel_l2(got, ref) <= tol
    else:
        assert np.max(np.abs(got - ref)) <= tol * np.max(np.abs(ref))
    assert np.array_equal(kfield_np(k), kin)            # k-field preserved
    assert np.all(got[:h] == 0) and np.all(got[:, :, -h:] == 0)  # real ghosts untouched


# ---------------------------------------------------------------- C2C PlanFFT
def _c2c_case(N, h, kh, C, prec, seed):
    r = lf.CField(lf.Lattice(3, N, h), C, DT[prec])
    k = lf.CField(lf.Lattice(3, N, kh), C, DT[prec])
    g = torch.Generator(device="cpu").manual_seed(seed)
    r.t.copy_(torch.randn(r.t.shape, generator=g, dtype=torch.float64).to(r.t.dtype))
    return r, k, lf.PlanFFT(r, k)


def _close(got, ref, prec, tol):
    if prec == "f32":
        return rel_l2(got, ref) <= tol
    return np.max(np.abs(got - ref)) <= tol * np.max(np.abs(ref))


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 2e-6)])
@pytest.mark.parametrize("N,h,kh,C", [(8, 1, 1, 1), (16, 0, 1, 1), (32, 2, 1, 3), (64, 1, 0, 1), (128, 1, 1, 1),
                                      (16, 1, 2, 6), (12, 1, 1, 1), (96, 1, 0, 2), (40, 1, 1, 1)])
def test_c2c_forward_matches_oracle(prec, tol, N, h, kh, C):
    """PlanFFT<Imag>(Field<Imag>*, Field<Imag>*) forward (LATfield2_PlanFFT_decl.hpp:1061-1133)"""
    announce(f"c2c forward {prec} N={N} h={h} kh={kh} C={C}")
    r, k, plan = _c2c_case(N, h, kh, C, prec, seed=3 * N + C)
    rin = kfield_np(r).copy()
    plan.execute(lf.FFT_FORWARD)
    torch.cuda.synchronize()
    ref = orc.fft_c2c_forward(rin, N, h, kh)
    got = kfield_np(k)
    assert _close(got, ref, prec, tol), rel_l2(got, ref)
    assert np.array_equal(kfield_np(r), rin)  # input preserved
    if kh:
        assert np.all(got[:kh] == 0) and np.all(got[:, :, :kh] == 0)


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 2e-6)])
@pytest.mark.parametrize("N,h,kh,C", [(8, 1, 0, 1), (32, 1, 1, 2), (64, 2, 1, 1), (256, 1, 1, 1), (48, 1, 1, 1)])
def test_c2c_backward_matches_oracle_inverse(prec, tol, N, h, kh, C):
    announce(f"c2c backward {prec} N={N} h={h} kh={kh} C={C}")
    r = lf.CField(lf.Lattice(3, N, h), C, DT[prec])
    k = lf.CField(lf.Lattice(3, N, kh), C, DT[prec])
    g = torch.Generator(device="cpu").manual_seed(N + 11 * C)
    k.t.copy_(torch.randn(k.t.shape, generator=g, dtype=torch.float64).to(k.t.dtype))
    plan = lf.PlanFFT(r, k)
    kin = kfield_np(k).copy()
    plan.execute(lf.FFT_BACKWARD)
    torch.cuda.synchronize()
    ref = orc.complex_field(N, h, C)
    orc.fft_c2c_backward(np.ascontiguousarray(kin), ref, N, h, kh)
    got = kfield_np(r)
    assert _close(got, ref, prec, tol), rel_l2(got, ref)
    assert np.array_equal(kfield_np(k), kin)  # k-field preserved
    if h:
        assert np.all(got[:h] == 0) and np.all(got[:, :, -h:] == 0)  # position ghosts untouched


@pytest.mark.parametrize("N", [512, 1024])
def test_c2c_full_size_round_trip_and_plane_wave_fp32(N):
    """size-independent properties at 512^3 / 1024^3: one plane wave -> N^3 at its cKSite,
    zeros elsewhere; backward(forward(f)) = N^3 f (both on the device)"""
    announce(f"c2c full size fp32 N={N}")
    lat = lf.Lattice(3, N, 1)
    r, k = lf.CField(lat, 1, lf.F32), lf.CField(lf.Lattice(3, N, 0), 1, lf.F32)
    plan = lf.PlanFFT(r, k)
    a, b, c = 3, N - 7, 5
    ar = torch.arange(N, device="cuda", dtype=torch.float64)
    ph = 2 * np.pi * (c * ar[:, None, None] + b * ar[None, :, None] + a * ar[None, None, :]) / N
    r.interior()[..., 0, 0] = torch.cos(ph).float()
    r.interior()[..., 0, 1] = torch.sin(ph).float()
    del ph
    plan.execute(lf.FFT_FORWARD)
    kc = k.as_complex()[..., 0]  # [ky, kx, kz]
    peak = kc[b, a, c].clone()
    kc[b, a, c] = 0
    off = torch.linalg.vector_norm(kc).item() / N ** 3
    assert abs(peak.real.item() / N ** 3 - 1) < 1e-5 and abs(peak.imag.item()) / N ** 3 < 1e-5
    assert off < 5e-6, off  # L2 of the other modes, relative to the peak
    g = torch.Generator(device="cuda").manual_seed(N)
    r.t.copy_(torch.randn(r.t.shape, generator=g, device="cuda", dtype=torch.float32))
    f0 = r.interior().clone()
    plan.execute(lf.FFT_FORWARD)
    plan.execute(lf.FFT_BACKWARD)
    torch.cuda.synchronize()
    err = torch.linalg.vector_norm(r.interior() / N ** 3 - f0) / torch.linalg.vector_norm(f0)
    assert err.item() <= 2e-6, err.item()


@pytest.mark.parametrize("c2c", [False, True])
def test_fp64_round_trips_at_1024(c2c):
    """R2C/C2R and C2C at 1024^3 in FP64 (the 128 KB strided tiles and the 144 KB C2C z tiles
    of the FP64 path): backward(forward(f)) / N^3 = f to 1e-12 relative (max over the mesh)"""
    N = 1024
    announce(f"fp64 round trip 1024 c2c={c2c}")
    lat = lf.Lattice(3, N, 1)
    g = torch.Generator(device="cuda").manual_seed(11)
    if c2c:
        r, k = lf.CField(lat, 1, lf.F64), lf.CField(lf.Lattice(3, N, 0), 1, lf.F64)
    else:
        r, k = lf.Field(lat, 1, lf.F64), lf.KField(lat, 0, 1, lf.F64)
    r.t.copy_(torch.randn(r.t.shape, generator=g, device="cuda", dtype=torch.float64))
    f0 = r.interior().clone()
    plan = lf.PlanFFT(r, k)
    plan.execute(lf.FFT_FORWARD)
    plan.execute(lf.FFT_BACKWARD)
    torch.cuda.synchronize()
    err = (r.interior() / N ** 3 - f0).abs().max().item() / f0.abs().max().item()
    assert err <= 1e-12, err
    del plan, r, k, f0
    torch.cuda.empty_cache()


@pytest.mark.parametrize("N", [384, 768, 320, 640])
@pytest.mark.parametrize("c2c", [False, True])
def test_round_trips_at_three_times_a_power_of_two(N, c2c):
    """R2C/C2R and C2C at 384^3, 768^3 (radix plans [8,4,4,3], [8,8,4,3]) and 320^3, 640^3
    ([8,8,5], [8,4,4,5]) in FP32:
    backward(forward(f)) / N^3 = f to 2e-6 rel-L2 on the device"""
    announce(f"round trip N={N} c2c={c2c}")
    lat = lf.Lattice(3, N, 1)
    g = torch.Generator(device="cuda").manual_seed(N)
    if c2c:
        r, k = lf.CField(lat, 1, lf.F32), lf.CField(lf.Lattice(3, N, 0), 1, lf.F32)
    else:
        r, k = lf.Field(lat, 1, lf.F32), lf.KField(lat, 0, 1, lf.F32)
    r.t.copy_(torch.randn(r.t.shape, generator=g, device="cuda", dtype=torch.float32))
    f0 = r.interior().clone()
    plan = lf.PlanFFT(r, k)
    plan.execute(lf.FFT_FORWARD)
    plan.execute(lf.FFT_BACKWARD)
    torch.cuda.synchronize()
    err = (torch.linalg.vector_norm(r.interior() / N ** 3 - f0) / torch.linalg.vector_norm(f0)).item()
    assert err <= 2e-6, err
    del plan, r, k, f0
    torch.cuda.empty_cache()


def test_fft_sizes_outside_the_plans_are_rejected():
    announce("fft size checks")
    for N in (6, 14, 36, 90):
        with pytest.raises(lf.Lf2miError):
            lf.PlanFFT(lf.Field(lf.Lattice(3, N, 1), 1, lf.F32), None)


def test_wave_test_kats_on_device_fp64():
    """benchmarks_tests/wave_test.cpp and wave_test_back.cpp (BoxSize 64, tolerance 1e-10)"""
    N, tol = 64, 1e-10
    announce("wave_test KATs fp64 N=64")
    lat = lf.Lattice(3, N, 1)
    r = lf.Field(lat, 1, lf.F64)
    k = lf.KField(lat, 0, 1, lf.F64)
    plan = lf.PlanFFT(r, k)
    x = np.arange(N)
    count = 0
    for i, j, l in [(0, 0, 0), (1, 2, 3), (3, 0, 2), (2, 3, 0), (1, 1, 1), (0, 3, 1)]:
        val = (np.cos(2 * np.pi * l * x / N)[:, None, None] * np.cos(2 * np.pi * j * x / N)[None, :, None]
               * np.cos(2 * np.pi * i * x / N)[None, None, :])
        r.interior()[..., 0] = torch.from_numpy(val).cuda()
        plan.execute(lf.FFT_FORWARD)
        got = kfield_np(k)[..., 0] / N ** 3  # [ky, kz, kx]
        amp = 1.0 / (2 ** sum(1 for q in (i, j, l) if q))
        exp = np.zeros_like(got)
        for ky in {j, (N - j) % N}:
            for kz in {l, (N - l) % N}:
                exp[ky, kz, i] = amp
        count += int(np.sum(np.abs(got - exp) > tol))
        # backward KAT: phiK = cos products in rKSite coordinates
        kv = (np.cos(2 * np.pi * j * x / N)[:, None, None] * np.cos(2 * np.pi * l * x / N)[None, :, None]
              * np.cos(2 * np.pi * i * np.arange(N // 2 + 1) / N)[None, None, :])
        k.t.zero_()
        k.t[..., 0, 0] = torch.from_numpy(kv).cuda()
        plan.execute(lf.FFT_BACKWARD)
        gr = real_np(r)[1:N + 1, 1:N + 1, 1:N + 1, 0] / N ** 3
        er = np.zeros_like(gr)
        for xx in {i, (N - i) % N}:
            for yy in {j, (N - j) % N}:
                for zz in {l, (N - l) % N}:
                    er[zz, yy, xx] = amp
        count += int(np.sum(np.abs(gr - er) > tol))
    assert count == 0


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 1e-6)])
def test_poisson_green_matches_oracle(prec, tol):
    N, kh = 32, 1
    announce(f"poisson_green {prec} N={N}")
    lat = lf.Lattice(3, N, 1)
    k = lf.KField(lat, kh, 2, DT[prec])
    g = torch.Generator(device="cpu").manual_seed(5)
    k.t.copy_(torch.randn(k.t.shape, generator=g, dtype=torch.float64).to(k.t.dtype))
    kin = np.ascontiguousarray(kfield_np(k))
    lf.poisson_green(k, res=0.5, scale=0.25)
    torch.cuda.synchronize()
    orc.poisson_green(kin, N, kh, res=0.5, scale=0.25)
    got = kfield_np(k)
    assert rel_l2(got, kin) <= tol


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 2e-6)])
@pytest.mark.parametrize("N", [16, 64, 128, 96, 80])
def test_fused_poisson_solve_matches_oracle_chain(prec, tol, N):
    """mean-subtract -> FFT -> -1/K^2 (zero mode 0) -> inverse FFT / N^3, as in orc_pm_step"""
    announce(f"poisson_solve {prec} N={N}")
    lat = lf.Lattice(3, N, 1)
    rho = _field(N, 1, 1, prec, seed=N)
    rho.t.add_(3.0)
    phi = lf.Field(lat, 1, DT[prec])
    plan = lf.PlanFFT(rho, None)
    rin = np.ascontiguousarray(real_np(rho))
    mean = rin[1:N + 1, 1:N + 1, 1:N + 1].mean()
    lf.poisson_solve(plan, rho, phi, shift=mean, scale=1.0 / N ** 3)
    torch.cuda.synchronize()
    d = rin.copy()
    d[1:N + 1, 1:N + 1, 1:N + 1] -= mean
    kk = orc.fft_r2c(d, N, 1, 0)
    orc.poisson_green(kk, N, 0, res=1.0 / N, scale=1.0 / N ** 3)
    ref = orc.real_field(N, 1, 1)
    orc.fft_c2r(kk, ref, N, 1, 0)
    got = real_np(phi)[1:N + 1, 1:N + 1, 1:N + 1]
    refi = ref[1:N + 1, 1:N + 1, 1:N + 1]
    if prec == "f32":
        assert rel_l2(got, refi) <= tol
    else:
        assert np.max(np.abs(got - refi)) <= tol * np.max(np.abs(refi))
    assert np.array_equal(real_np(rho), rin)  # rho is read only


@pytest.mark.parametrize("prec,tol", [("f32", 1e-6), ("f64", 1e-12)])
@pytest.mark.parametrize("N", [512, 1024])
def test_fused_poisson_solve_large_matches_unfused_chain(N, prec, tol):
    """the fused solve at the benchmark sizes (1024^3 runs its workspace transposed; the
    y + Green pass's tiles are remapped over the XCDs where their rows are under 128 bytes:
    FP32 at N >= 512, FP64 at 1024) against execute(FFT_FORWARD) -> poisson_green ->
    execute(FFT_BACKWARD), whose kernels are pinned against the oracle at smaller N and by the
    round trips at 1024^3; rel-L2 <= 1e-6 (FP32), 1e-12 (FP64)"""
    esz = 4 if prec == "f32" else 8
    free, _ = torch.cuda.mem_get_info()
    need = 8 * (N + 2) ** 3 * esz
    if free < need:
        pytest.skip(f"needs {need / 2 ** 30:.0f} GiB of device memory")
    announce(f"poisson_solve fused vs unfused {prec} N={N}")
    dt = DT[prec]
    lat = lf.Lattice(3, N, 1)
    rho = lf.Field(lat, 1, dt)
    g = torch.Generator(device="cuda").manual_seed(N)
    tdt = torch.float32 if prec == "f32" else torch.float64
    rho.t.copy_(torch.rand(rho.t.shape, generator=g, device="cuda", dtype=tdt) + 2.0)
    inner = rho.t[1:N + 1, 1:N + 1, 1:N + 1, 0]
    mean = float(inner.double().mean().item())
    phi = lf.Field(lat, 1, dt)
    plan = lf.PlanFFT(rho, None)
    lf.poisson_solve(plan, rho, phi, shift=mean, scale=1.0 / N ** 3)
    del plan
    # the unfused chain on the mean-subtracted field (the same subtraction)
    inner -= mean
    k = lf.KField(lat, 0, 1, dt)
    plan2 = lf.PlanFFT(rho, k)
    plan2.execute(lf.FFT_FORWARD)
    lf.poisson_green(k, res=1.0 / N, scale=1.0 / N ** 3)
    ref = lf.Field(lat, 1, dt)
    plan3 = lf.PlanFFT(ref, k)
    plan3.execute(lf.FFT_BACKWARD)
    torch.cuda.synchronize()
    a = phi.t[1:N + 1, 1:N + 1, 1:N + 1, 0].double()
    b = ref.t[1:N + 1, 1:N + 1, 1:N + 1, 0].double()
    err = float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))
    print(f"N={N} {prec}: fused vs unfused rel-L2 {err:.2e}")
    assert err <= tol
