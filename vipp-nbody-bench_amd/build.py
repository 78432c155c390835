"""Build recipe of libvipp_gravity.so (hand-written HIP for gfx950).

    python vipp-nbody-bench_amd/build.py        # from the repository root

Every source under csrc/ is compiled by hipcc for gfx950 only, with packed
FP32 instructions disabled, and linked into lib/libvipp_gravity.so, which the
C++ wrapper, the tests and bench.py load.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "lib", "libvipp_gravity.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
SOURCES = ["api.hip", "tree.hip", "sort.hip", "fmm.hip", "p2p.hip", "rk4.hip", "m2l_mfma.hip", "direct.hip", "mlp.hip", "closure.hip"]
# host-only sources, compiled by the host compiler the reference is built with
# (g++ without FMA contraction): problem.cpp must reproduce its libm calls
HOST_SOURCES = ["problem.cpp"]
FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function", "-munsafe-fp-atomics"]


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, src.replace(".hip", ".o"))
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr}")
    return obj


def _compile_host(src: str) -> str:
    obj = os.path.join(OBJ, src.replace(".cpp", ".o"))
    cmd = [os.environ.get("CXX", "g++"), "-std=c++17", "-O2", "-fPIC", "-ffp-contract=off", "-Wall", "-Wextra", "-I",
           os.path.join(os.path.dirname(HERE), "include"), "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"host compile failed on {src}:\n{r.stderr}")
    return obj


def build(jobs: int = 8) -> str:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES)) + list(ex.map(_compile_host, HOST_SOURCES))
    tmp = LIB + ".tmp"
    r = subprocess.run([HIPCC, "-shared", "--offload-arch=gfx950", "-o", tmp, *objs], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    build_cpp_test()
    return LIB


ROOT = os.path.dirname(HERE)
CPP_TEST = os.path.join(OBJ, "solver_gpu_test")


def build_cpp_test() -> str:
    """The C++ wrapper's test program (tests/cpp/solver_gpu_test.cpp), host code
    linked against the library."""
    src = os.path.join(ROOT, "tests", "cpp", "solver_gpu_test.cpp")
    cmd = [os.environ.get("CXX", "g++"), "-std=c++17", "-O2", "-Wall", "-Wextra", "-I", os.path.join(ROOT, "include"), src,
           "-L", os.path.dirname(LIB), "-lvipp_gravity", "-Wl,-rpath,$ORIGIN/../lib", "-o", CPP_TEST]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"C++ wrapper test build failed:\n{r.stderr}")
    return CPP_TEST


if __name__ == "__main__":
    print(build(int(sys.argv[1]) if len(sys.argv) > 1 else 8))
