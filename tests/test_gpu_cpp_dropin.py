"""GPU: the C++ drop-in mirror (include/LATfield2_mi.hpp) driven the way a LATfield2
user's program drives the reference.  The programs in tests/cpp/ restate the reference's
known-answer programs (benchmarks_tests/wave_test.cpp, wave_test_back.cpp,
examples/poissonSolver.cpp, examples/fft_c2c_r2c_sametime.cpp, examples/boundary_test.cpp) and a
projection/particle check; each exits with the number
of failures.  Built by `python -m latfield2_amd.build` (and __graft_entry__.build)."""
import os
import subprocess

import pytest

from gpu_util import announce

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "latfield2_amd", "_bin")


def _bin(name):
    path = os.path.join(BIN, name)
    if not os.path.exists(path):
        from latfield2_amd import build
        build.build_cpp_tests(verbose=False)
    return path


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("prog", ["wave_test_mi", "wave_test_back_mi", "poisson_solver_mi", "particles_mi",
                                  "fft_c2c_mi", "boundary_test_mi"])
def test_cpp_dropin_program(prog, prec):
    exe = _bin(f"{prog}_{prec}")
    announce(f"{prog}_{prec}")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=110)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
