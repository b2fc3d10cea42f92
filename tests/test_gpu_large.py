"""Large multi-component projection on one GPU: 1024^3 mesh with 1024^3 particles (about
26 GB of FP32 particle state, 60 GB with the re-bucketing buffer), positions on the lattice
plus a Gaussian displacement (sigma 1 cell), velocities Gaussian (sigma 0.01), mass 1/Np.
Per kind: the scalar mass deposit, the 3-component vectorProjectionCIC (m v_i) and the
6-component projection_Tij (m v_i v_j), each followed by R2C forward and C2R backward FFTs
of every component; and the two CIC-NGP hybrids (deposit only).

Parity: the FP64 oracle deposits the particles of a 256^3 window (cells [x0-1, x0+256] per
axis, downloaded with lf2mi_parts_download_box) into a 512^3 periodic box shifted by
x0 - 2 cells; a shift by whole cells keeps every particle's cell offset, and nothing wraps
in that box, so its nodes [2, 258) equal the 1024^3 field's nodes [x0, x0+256) up to the
density factor (1024/512)^3.  Bars: deposits rel-L2 <= 1e-6 on the window; the FFT round
trip (C2R(R2C(f))/N^3) rel-L2 <= 2e-6 and Parseval's identity to 1e-5 on the whole mesh.
The kernel times of the deposits and FFT passes are printed with their HBM GB/s.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import latfield2_amd as lf  # noqa: E402
import oracle as orc  # noqa: E402
from latfield2_amd import _lib  # noqa: E402
from gpu_util import announce, rel_l2  # noqa: E402

pytestmark = pytest.mark.gpu

N = 1024
W = 256
X0 = (136, 392, 648)   # window origin (cells), not aligned to bricks
NP = 512               # oracle box


def _sumsq_real(t, C):
    """sum of squares of the interior of a (G, G, G, C) field, in FP64, in z slabs"""
    tot = 0.0
    for z0 in range(1, N + 1, 64):
        tot += float(t[z0:min(z0 + 64, N + 1), 1:N + 1, 1:N + 1, :C].double().pow(2).sum().item())
    return tot


def _sumsq_modes(kc, C):
    """sum over the full spectrum of |F|^2 from the R2C half (kx = 1..N/2-1 count twice)"""
    w = torch.full((N // 2 + 1,), 2.0, dtype=torch.float64, device="cuda")
    w[0] = 1.0
    w[N // 2] = 1.0
    tot = 0.0
    for c in range(C):
        for y0 in range(0, N, 64):
            a = kc[y0:y0 + 64, :, :, c].abs().double().pow(2)
            tot += float((a * w).sum().item())
    return tot


def _profile(fn):
    torch.cuda.synchronize()
    _lib.profile_reset()
    _lib.profile_enable(True)
    fn()
    torch.cuda.synchronize()
    _lib.profile_enable(False)
    return {k: t / c for k, (t, c) in _lib.profile_read().items()}


@pytest.fixture(scope="module")
def big():
    free, _ = torch.cuda.mem_get_info()
    if free < 180 * 2 ** 30:
        pytest.skip(f"needs about 180 GiB of free device memory, {free / 2 ** 30:.0f} GiB free")
    announce(f"large projection {N}^3, {N}^3 particles")
    parts = lf.Particles(N, N ** 3, lf.F32, mass=1.0 / N ** 3)
    parts.ic_lattice(N, 1.0, 0.01, 5)
    lo = [x - 1 for x in X0]
    hi = [x + W + 1 for x in X0]
    box = parts.download_box(lo, hi, 24_000_000)
    # the window's particles in the shifted oracle box (exact: power-of-two scalings and
    # whole-cell shifts)
    p = box["pos"] * N - (np.array(X0, dtype=np.float64) - 2.0)
    pos = np.ascontiguousarray(p / NP)
    assert pos.min() > 0.0 and pos.max() < (W + 4) / NP
    yield parts, pos, np.ascontiguousarray(box["vel"]), np.ascontiguousarray(box["mass"])
    del parts
    torch.cuda.empty_cache()


@pytest.mark.parametrize("kind,C", [(lf.PROJ_SCALAR, 1), (lf.PROJ_VECTOR, 3), (lf.PROJ_VECVEC, 6)])
def test_large_projection_and_ffts(big, kind, C):
    parts, pos, vel, mass = big
    name = {lf.PROJ_SCALAR: "scalar", lf.PROJ_VECTOR: "vector", lf.PROJ_VECVEC: "VecVec"}[kind]
    announce(f"large projection {name}, C={C}")
    lat = lf.Lattice(3, N, 1)
    f = lf.Field(lat, C, lf.F32)
    f.zero()
    t_dep = _profile(lambda: parts.deposit(kind, f, lf.DEP_PERIODIC))
    sl = tuple(slice(1 + x, 1 + x + W) for x in (X0[2], X0[1], X0[0]))  # [z, y, x]
    got = f.t[sl].double().cpu().numpy()
    # oracle on the window
    orc.set_threads(16)
    ref = orc.real_field(NP, 1, C)
    if kind == lf.PROJ_SCALAR:
        orc.scalar_project(pos, ref, NP, 1, mass=mass)
    elif kind == lf.PROJ_VECTOR:
        orc.vector_project(pos, vel, ref, NP, 1, mass=mass)
    else:
        orc.vecvec_project(pos, vel, ref, NP, 1, mass=mass)
    ref = ref[3:3 + W, 3:3 + W, 3:3 + W] * (N / NP) ** 3   # nodes [2, 258) with the ghost offset 1
    err = rel_l2(got, ref)
    # FFTs of every component: forward, backward, round trip and Parseval
    k = lf.KField(lat, 0, C, lf.F32)
    plan = lf.PlanFFT(f, k)
    e2 = _sumsq_real(f.t, C)
    t_fwd = _profile(lambda: plan.execute(lf.FFT_FORWARD))
    kc = k.as_complex()
    ek = _sumsq_modes(kc, C)
    t_bwd = _profile(lambda: plan.execute(lf.FFT_BACKWARD))
    back = f.t[sl].double().cpu().numpy() / N ** 3
    e_rt = rel_l2(back, got)
    e_pars = abs(ek / N ** 3 - e2) / e2
    M = (N // 2 + 1) * N * N
    dep_bytes = {1: 16, 3: 28, 6: 28}[C] * N ** 3 + 4 * C * N ** 3
    fwd_bytes = C * (4 * N ** 3 + 8 * M + 2 * 16 * M)
    ms_dep = sum(t_dep.values())
    ms_fwd = sum(t_fwd.values())
    ms_bwd = sum(t_bwd.values())
    print(f"{name}: deposit {ms_dep:.2f} ms ({dep_bytes / ms_dep / 1e6:.0f} GB/s) {t_dep}; "
          f"FFT forward {ms_fwd:.2f} ms ({fwd_bytes / ms_fwd / 1e6:.0f} GB/s) {t_fwd}; "
          f"backward {ms_bwd:.2f} ms ({fwd_bytes / ms_bwd / 1e6:.0f} GB/s) {t_bwd}; "
          f"window deposit rel-L2 {err:.2e}, round trip {e_rt:.2e}, Parseval {e_pars:.2e}")
    assert err <= 1e-6
    assert e_rt <= 2e-6
    assert e_pars <= 1e-5
    del f, k, plan, kc
    torch.cuda.empty_cache()


@pytest.mark.parametrize("kind,C", [(lf.PROJ_VECTOR_NGP, 3), (lf.PROJ_SYMTENSOR_NGP, 6)])
def test_large_cic_ngp_projections(big, kind, C):
    """vectorProjectionCICNGP / symtensorProjectionCICNGP at 1024^3 (projections.hpp:405-976):
    the same window parity as the CIC kinds"""
    parts, pos, vel, mass = big
    name = {lf.PROJ_VECTOR_NGP: "vector CIC-NGP", lf.PROJ_SYMTENSOR_NGP: "symtensor CIC-NGP"}[kind]
    announce(f"large projection {name}, C={C}")
    lat = lf.Lattice(3, N, 1)
    f = lf.Field(lat, C, lf.F32)
    f.zero()
    t_dep = _profile(lambda: parts.deposit(kind, f, lf.DEP_PERIODIC))
    sl = tuple(slice(1 + x, 1 + x + W) for x in (X0[2], X0[1], X0[0]))  # [z, y, x]
    got = f.t[sl].double().cpu().numpy()
    orc.set_threads(16)
    ref = orc.real_field(NP, 1, C)
    if kind == lf.PROJ_VECTOR_NGP:
        orc.vector_ngp_project(pos, vel, ref, NP, 1, mass=mass)
    else:
        orc.symtensor_ngp_project(pos, vel, ref, NP, 1, mass=mass)
    ref = ref[3:3 + W, 3:3 + W, 3:3 + W] * (N / NP) ** 3
    err = rel_l2(got, ref)
    ms_dep = sum(t_dep.values())
    dep_bytes = (28 + 4 * C) * N ** 3
    print(f"{name}: deposit {ms_dep:.2f} ms ({dep_bytes / ms_dep / 1e6:.0f} GB/s) {t_dep}; "
          f"window deposit rel-L2 {err:.2e}")
    assert err <= 1e-6
    del f
    torch.cuda.empty_cache()
