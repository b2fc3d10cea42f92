"""CPU checks of the C++ drop-in mirror: the header compiles with the reference-style
programs (syntax only) and its host-side geometry (Site, rKSite, symmetric fields)
is right.  No GPU work."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CXX = shutil.which("g++")


@pytest.mark.skipif(CXX is None, reason="no host C++ compiler")
def test_shim_host_geometry(tmp_path):
    exe = tmp_path / "shim_host_check"
    src = os.path.join(ROOT, "tests", "cpp_host", "shim_host_check.cpp")
    # only header code runs: link against the library for the unreferenced ABI symbols
    subprocess.run([CXX, "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"), src, "-o", str(exe),
                    "-L", os.path.join(ROOT, "latfield2_amd"), "-llf2mi",
                    "-Wl,-rpath," + os.path.join(ROOT, "latfield2_amd")], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.skipif(CXX is None, reason="no host C++ compiler")
@pytest.mark.parametrize("prog", sorted(f for f in os.listdir(os.path.join(ROOT, "tests", "cpp")) if f.endswith(".cpp")))
@pytest.mark.parametrize("defs", [[], ["-DSINGLE"]])
def test_dropin_programs_compile(prog, defs):
    subprocess.run([CXX, "-fsyntax-only", "-std=c++17", "-Wall", "-Werror", *defs, "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", prog)], check=True)


REFERENCE = "/root/reference"
REFERENCE_DRIVERS = ["examples/poissonSolver.cpp", "benchmarks_tests/wave_test.cpp",
                     "benchmarks_tests/wave_test_back.cpp", "examples/fft_c2c_r2c_sametime.cpp",
                     "examples/fft.cpp", "examples/gettingStarted.cpp", "examples/test_sum.cpp"]


@pytest.mark.skipif(CXX is None, reason="no host C++ compiler")
@pytest.mark.skipif(not os.path.isdir(REFERENCE), reason="the reference tree is not present")
@pytest.mark.parametrize("driver", REFERENCE_DRIVERS)
@pytest.mark.parametrize("defs", [[], ["-DSINGLE"]])
def test_reference_drivers_compile_against_the_dropin_header(driver, defs):
    """the reference's own programs, read where they lie, compile unmodified against
    include/LATfield2.hpp (parallel, COUT, Lattice, Site/rKSite/cKSite, Field, PlanFFT,
    MPI_Wtime as they use them)"""
    subprocess.run([CXX, "-fsyntax-only", "-std=c++17", *defs, "-I", os.path.join(ROOT, "include"),
                    os.path.join(REFERENCE, driver)], check=True)
