"""Builds liblf2mi.so (the HIP kernels + C ABI) in-tree for gfx950.

    python -m latfield2_amd.build          # compile every source, link the library

Objects go to latfield2_amd/_build/, the library to latfield2_amd/liblf2mi.so.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "liblf2mi.so")
SOURCES = ["field.hip", "fft.hip", "particles.hip", "diffpm.hip"]
BIN = os.path.join(PKG, "_bin")
CPP_TESTS = os.path.join(ROOT, "tests", "cpp")
CXX = os.environ.get("CXX_HOST", shutil.which("g++") or "g++")

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-munsafe-fp-atomics",
    # keep packed FP32 instructions out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-value", "-Wno-unused-result",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    jobs = min(len(SOURCES), int(os.environ.get("MAX_JOBS", "8")))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


def build_cpp_tests(verbose: bool = True) -> list:
    """C++ programs written against include/LATfield2_mi.hpp (the drop-in mirror of the
    reference's API), each in double and in SINGLE precision, linked to liblf2mi.so."""
    os.makedirs(BIN, exist_ok=True)
    jobs = []
    for src in sorted(f for f in os.listdir(CPP_TESTS) if f.endswith(".cpp")):
        base = os.path.splitext(src)[0]
        for tag, defs in (("f64", []), ("f32", ["-DSINGLE"])):
            out = os.path.join(BIN, f"{base}_{tag}")
            jobs.append([CXX, "-O2", "-std=c++17", "-Wall", *defs, "-I", os.path.join(ROOT, "include"),
                         os.path.join(CPP_TESTS, src), "-L", PKG, "-llf2mi", "-Wl,-rpath,$ORIGIN/..", "-o", out])

    def run(cmd):
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"C++ build failed: {' '.join(cmd)}\n{r.stderr[-4000:]}")
        return cmd[-1]

    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        outs = list(ex.map(run, jobs))
    if verbose:
        print(f"built {len(outs)} C++ drop-in programs in {BIN}")
    return outs


if __name__ == "__main__":
    try:
        build()
        build_cpp_tests()
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
