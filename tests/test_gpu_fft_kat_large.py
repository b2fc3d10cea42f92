"""GPU: the known-answer checks of the reference's FFT tests at the benchmark sizes, where the
transforms run their largest plans (512^3: radix 8 8 8; 1024^3: 8 8 4 4, with the fused
Poisson workspace transposed), so that a layout, sign or normalisation error of those plans
cannot pass unseen.

* benchmarks_tests/wave_test.cpp:72-152: phi = cos(2 pi i x/N) cos(2 pi j y/N) cos(2 pi l z/N)
  transforms to amp N^3 at kx = i, ky = +-j, kz = +-l (rKSite coordinates, k-field stored
  (kx, kz, ky)) and to 0 elsewhere; amp = 1/2 per factor whose index is neither 0 nor N/2.
* benchmarks_tests/wave_test_back.cpp:76-138: the same cos products given in k space come
  back as amp N^3 at x = +-i, y = +-j, z = +-l.
* examples/poissonSolver.cpp:99-111 in closed form: rho = cos(2 pi (a x + b y + c z)/N) has
  phi = -rho / K^2 with K^2 = sum_d N^2 (2 - 2 cos(2 pi k_d / N)) (the solve's 1/N^3
  included), fused and through execute + poisson_green + execute.
The expected fields are formed and compared on the device (no oracle run at these sizes).
Bars: FP32 |error| / N^3 <= 1e-6 per mode and site, Poisson max error <= 2e-6 of max |phi|;
FP64 1e-10 (the reference's wave_test bar) and 1e-11.
"""
import math

import pytest

torch = pytest.importorskip("torch")

import latfield2_amd as lf  # noqa: E402
from gpu_util import announce  # noqa: E402

pytestmark = pytest.mark.gpu
DT = {"f32": lf.F32, "f64": lf.F64}
TD = {"f32": torch.float32, "f64": torch.float64}


def _need(nbytes):
    free, _ = torch.cuda.mem_get_info()
    if free < nbytes:
        pytest.skip(f"needs {nbytes / 2 ** 30:.0f} GiB of device memory")


def _amp(q, N):
    return 1.0 if q in (0, N // 2) else 0.5


def _cos(q, n, N):
    x = torch.arange(n, device="cuda", dtype=torch.int64)
    return torch.cos(2.0 * math.pi * ((q * x) % N).to(torch.float64) / N)


def _modes(N):
    return [(0, 0, 0), (1, 2, 3), (N // 2, 5, 7), (3, N // 2, 1), (7, 11, N // 2), (N // 2 - 1, N - 3, 13)]


@pytest.mark.parametrize("N,prec,tol", [(512, "f32", 1e-6), (1024, "f32", 1e-6), (512, "f64", 1e-10)])
def test_wave_test_kats_at_benchmark_sizes(N, prec, tol):
    esz = 4 if prec == "f32" else 8
    _need(esz * (N + 2) ** 3 * 5 + 8 * N ** 3 * 2)
    lat = lf.Lattice(3, N, 1)
    r = lf.Field(lat, 1, DT[prec])
    k = lf.KField(lat, 0, 1, DT[prec])
    plan = lf.PlanFFT(r, k)
    n3 = float(N) ** 3
    worst_f = worst_b = 0.0
    for (i, j, l) in _modes(N):
        announce(f"wave_test KAT {prec} N={N} mode=({i},{j},{l})")
        amp = _amp(i, N) * _amp(j, N) * _amp(l, N)
        # forward: real cos products -> one amplitude per (+-j, +-l)
        r.interior()[..., 0] = (_cos(l, N, N)[:, None, None] * _cos(j, N, N)[None, :, None]
                                * _cos(i, N, N)[None, None, :]).to(TD[prec])
        plan.execute(lf.FFT_FORWARD)
        kc = k.as_complex()[..., 0]  # [ky, kz, kx]
        for ky in {j % N, (N - j) % N}:
            for kz in {l % N, (N - l) % N}:
                kc[ky, kz, i] -= amp * n3
        err_f = float(kc.abs().max()) / n3
        # backward: the cos products as k-space values (rKSite coordinates)
        k.t.zero_()
        k.t[..., 0, 0] = (_cos(j, N, N)[:, None, None] * _cos(l, N, N)[None, :, None]
                          * _cos(i, N // 2 + 1, N)[None, None, :]).to(TD[prec])
        plan.execute(lf.FFT_BACKWARD)
        ri = r.interior()[..., 0]  # [z, y, x]
        for xx in {i % N, (N - i) % N}:
            for yy in {j % N, (N - j) % N}:
                for zz in {l % N, (N - l) % N}:
                    ri[zz, yy, xx] -= amp * n3
        err_b = float(ri.abs().max()) / n3
        print(f"N={N} {prec} mode ({i},{j},{l}): forward {err_f:.2e}, backward {err_b:.2e}")
        worst_f, worst_b = max(worst_f, err_f), max(worst_b, err_b)
    assert worst_f <= tol and worst_b <= tol, (worst_f, worst_b)


@pytest.mark.parametrize("N,prec,tol", [(1024, "f32", 2e-6), (1024, "f64", 1e-11), (512, "f64", 1e-11)])
def test_poisson_closed_form_at_benchmark_sizes(N, prec, tol):
    """a single cosine mode through the fused solve and through execute + poisson_green +
    execute; the modes reach the Nyquist plane and the far end of each axis"""
    esz = 4 if prec == "f32" else 8
    _need(esz * (N + 2) ** 3 * 6 + 8 * N ** 3 * 2)
    lat = lf.Lattice(3, N, 1)
    rho = lf.Field(lat, 1, DT[prec])
    phi = lf.Field(lat, 1, DT[prec])
    k = lf.KField(lat, 0, 1, DT[prec])
    plan = lf.PlanFFT(rho, k)
    phi_plan = lf.PlanFFT(phi, k)
    x = torch.arange(N, device="cuda", dtype=torch.int64)
    worst = 0.0
    for (a, b, c) in [(1, 0, 0), (3, N - 7, N // 2), (N // 2, 1, N - 1), (17, 29, 5)]:
        announce(f"poisson closed form {prec} N={N} mode=({a},{b},{c})")
        ph = ((a * x[None, None, :] + b * x[None, :, None] + c * x[:, None, None]) % N).to(torch.float64)
        want = torch.cos(2.0 * math.pi * ph / N)
        del ph
        # 2 - 2 cos(x) = 4 sin^2(x / 2), the form without cancellation at small k
        K2 = sum(N * N * 4.0 * math.sin(math.pi * q / N) ** 2 for q in (a, b, c))
        rho.interior()[..., 0] = want.to(TD[prec])
        want.mul_(-1.0 / K2)
        wmax = float(want.abs().max())
        # fused
        lf.poisson_solve(plan, rho, phi, shift=0.0, scale=1.0 / float(N) ** 3)
        e1 = float((phi.interior()[..., 0].double() - want).abs().max()) / wmax
        # unfused: execute(FORWARD), Green (res = 1/N, lambda = -K^2), execute(BACKWARD)
        plan.execute(lf.FFT_FORWARD)
        lf.poisson_green(k, res=1.0 / N, scale=1.0 / float(N) ** 3)
        phi.t.zero_()
        phi_plan.execute(lf.FFT_BACKWARD)
        e2 = float((phi.interior()[..., 0].double() - want).abs().max()) / wmax
        print(f"N={N} {prec} mode ({a},{b},{c}): fused {e1:.2e}, unfused {e2:.2e}")
        worst = max(worst, e1, e2)
        del want
    assert worst <= tol, worst
