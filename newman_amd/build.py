"""Build the native library newman_amd/libnewman_amd.so (HIP kernels for gfx950
plus the C++ host side and the C ABI) and, when asked, the CPU oracle.

Usage: python -m newman_amd.build [--oracle]

The HIP file is compiled with -ffp-contract=off (the plain path must reproduce the
reference's FMA-free double arithmetic bit for bit) and with packed FP32 disabled.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libnewman_amd.so")
# test support: the host-side walk of the kernels' indexing (csrc/index_check.cpp)
INDEX_CHECK_LIB = os.path.join(HERE, "libnm_index_check.so")
ARCH = os.environ.get("NEWMAN_AMD_ARCH", "gfx950")

HOST_SRCS = ["hpfloat.cpp", "mandelbrot.cpp", "palette.cpp", "video.cpp", "capi.cpp"]
HIP_SRCS = ["kernels.hip", "skipnet.hip", "colour.hip", "video.hip"]

CXXFLAGS = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-Wall", "-Wno-unused-function",
            "-pthread"]
HIPFLAGS = ["--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-result"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build_library(verbose=True):
    """Every source compiled in parallel, then linked."""
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    for src in HOST_SRCS:
        jobs.append((["g++"] + CXXFLAGS + ["-c", os.path.join(CSRC, src), "-o", os.path.join(BUILD, src + ".o")]))
    for src in HIP_SRCS:
        jobs.append((["hipcc"] + HIPFLAGS + ["-c", os.path.join(CSRC, src), "-o", os.path.join(BUILD, src + ".o")]))
    procs = []
    for cmd in jobs:
        print(" ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
    failed = [cmd for cmd, p in zip(jobs, procs) if p.wait() != 0]
    if failed:
        raise subprocess.CalledProcessError(1, failed[0])
    objs = [cmd[-1] for cmd in jobs]
    tmp = LIB + ".tmp"
    _run(["hipcc", "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"])
    os.replace(tmp, LIB)
    return LIB


def build_index_check():
    """The indexing checker: host code over csrc/sample_index.hpp, which the kernels
    share; the HIP headers supply its host/device qualifiers."""
    tmp = INDEX_CHECK_LIB + ".tmp"
    _run(["g++"] + CXXFLAGS + ["-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-shared", "-o", tmp,
                               os.path.join(CSRC, "index_check.cpp")])
    os.replace(tmp, INDEX_CHECK_LIB)
    return INDEX_CHECK_LIB


def build_oracle():
    """The CPU oracle is test infrastructure (oracle/); built here for convenience."""
    _run(["make", "-B", "-C", os.path.join(ROOT, "oracle")])  # always recompiled, like the library
    return os.path.join(ROOT, "oracle", "liboracle.so")


if __name__ == "__main__":
    build_library()
    build_index_check()
    if "--oracle" in sys.argv:
        build_oracle()
