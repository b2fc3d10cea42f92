"""GPU: the C++ drop-in mirror (include/LATfield2.hpp -> LATfield2_mi.hpp) driven the way a
LATfield2 user's program drives the reference.

* The programs in tests/cpp/ restate the reference's known-answer programs
  (benchmarks_tests/wave_test.cpp, wave_test_back.cpp, examples/poissonSolver.cpp with its
  parallel/COUT calls and no explicit transfers, examples/fft_c2c_r2c_sametime.cpp,
  examples/boundary_test.cpp) and check the mirror's Field/Site/PlanFFT semantics
  (field_api_mi.cpp) and particles (particles_mi.cpp); each exits with the number of
  failures.  Built by `python -m latfield2_amd.build` (and __graft_entry__.build).
* The reference's own driver programs, compiled unmodified against the drop-in header where
  the reference tree was present (oracle/Makefile `dropin`, into oracle/_ref/), run as they
  are: their exit codes and printed checks are the reference's known answers."""
import os
import re
import subprocess

import pytest

from gpu_util import announce

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "latfield2_amd", "_bin")
REF_BIN = os.path.join(ROOT, "oracle", "_ref")


def _bin(name):
    path = os.path.join(BIN, name)
    if not os.path.exists(path):
        from latfield2_amd import build
        build.build_cpp_tests(verbose=False)
    return path


def _run(exe, args=(), cwd=None, timeout=110):
    r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=timeout, cwd=cwd)
    print(r.stdout[-3000:], r.stderr[-2000:])
    return r


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("prog", ["wave_test_mi", "wave_test_back_mi", "poisson_solver_mi", "particles_mi",
                                  "fft_c2c_mi", "boundary_test_mi", "field_api_mi"])
def test_cpp_dropin_program(prog, prec, tmp_path):
    exe = _bin(f"{prog}_{prec}")
    announce(f"{prog}_{prec}")
    r = _run(exe, cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_functor_without_device_implementation_is_refused(prec, tmp_path):
    """updateVel with a user's host functor: message and exit(-1), as the reference aborts"""
    exe = _bin(f"functor_refuse_mi_{prec}")
    announce(f"functor_refuse_mi_{prec}")
    r = _run(exe, cwd=str(tmp_path))
    assert r.returncode == 255 and "no device implementation" in r.stderr and "not reached" not in r.stdout


def _ref_bin(name):
    path = os.path.join(REF_BIN, name)
    if not os.path.exists(path):
        pytest.skip("the reference's drivers are built only where the reference tree is present "
                    "(make -C oracle dropin)")
    return path


@pytest.mark.parametrize("prog", ["poissonSolver_f64", "wave_test_f64", "wave_test_back_f64", "fft_f64"])
def test_reference_driver_unmodified(prog, tmp_path):
    """examples/poissonSolver.cpp exits 0 when its maximum error is within 1e-11
    (poissonSolver.cpp:153-159); wave_test.cpp / wave_test_back.cpp exit with the number
    of modes beyond 1e-10 (wave_test.cpp:119-172) over 64 cos-product inputs at 64^3;
    examples/fft.cpp runs forward and backward R2C and C2C plans."""
    exe = _ref_bin(prog)
    announce(f"reference driver {prog}")
    r = _run(exe, ["-n", "1", "-m", "1"], cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    if prog == "poissonSolver_f64":
        m = re.search(r"Max error: ([0-9.eE+-]+)", r.stdout)
        assert m and float(m.group(1)) <= 1e-11


def test_reference_fft_c2c_r2c_sametime(tmp_path):
    """examples/fft_c2c_r2c_sametime.cpp: a constant field through R2C forward + backward
    averages N^3 (printed as 1 after its 1/N^6); through C2C the reference's single-rank
    backward skips a transposition and its example prints 1 after dividing by N^5, while
    this build's C2C backward is the inverse of the forward, so the same line prints N = 16,
    the example's lattice size (the deliberate difference of the C2C backward)."""
    exe = _ref_bin("fft_c2c_r2c_sametime_f64")
    announce("reference driver fft_c2c_r2c_sametime_f64")
    r = _run(exe, ["-n", "1", "-m", "1"], cwd=str(tmp_path))
    assert r.returncode == 0
    reals = [float(v) for v in re.findall(r"Average real = ([0-9.eE+-]+)", r.stdout)]
    comp = re.search(r"Average complex = ([0-9.eE+-]+)", r.stdout)
    assert len(reals) == 2 and all(abs(v - 1.0) < 1e-12 for v in reals)
    assert comp and abs(float(comp.group(1)) - 16.0) < 1e-9


def test_reference_getting_started(tmp_path):
    """examples/gettingStarted.cpp on its 25 x 57 x 32 lattice: updateHalo, host stencils and
    Field::write of phi (the interior, 25*57*32 doubles, x fastest; phi = 1 + x)"""
    import numpy as np
    exe = _ref_bin("gettingStarted_f64")
    announce("reference driver gettingStarted_f64")
    r = _run(exe, ["-n", "1", "-m", "1"], cwd=str(tmp_path))
    assert r.returncode == 0
    assert "Lattice size: (25,57,32)" in r.stdout
    data = np.fromfile(os.path.join(str(tmp_path), "test_phi.txt"), dtype=np.float64)
    assert data.size == 25 * 57 * 32
    want = np.broadcast_to(1.0 + np.arange(25, dtype=np.float64), (32, 57, 25)).ravel()
    assert np.array_equal(data, want)
