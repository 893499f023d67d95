"""Build the HIP library ``mycode_amd/libcfdgpu.so`` for gfx950.

Host-only C++ (the mesh pipeline) is compiled with g++; the kernels and the
C-ABI with hipcc.  Multiply-add contraction is off everywhere so that the FP64
residual reproduces the reference's arithmetic exactly; packed FP32 ops are
disabled on the device.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libcfdgpu.so")
BUILD = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CFD_OFFLOAD_ARCH", "gfx950")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off"]
HIP_FLAGS = COMMON + [
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return r


ROOT = os.path.dirname(HERE)
HOST_TEST_SRC = os.path.join(ROOT, "tests", "host_abi_test.cpp")
HOST_TEST_BIN = os.path.join(ROOT, "tests", "host_abi_test")


def build_host_test(verbose: bool = False) -> str:
    """Compile tests/host_abi_test.cpp, a C++ host that uses only
    include/cfd_gpu.h, against the library (run by the GPU tests)."""
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror", f"-I{os.path.join(ROOT, 'include')}",
           HOST_TEST_SRC, f"-L{HERE}", "-lcfdgpu", "-Wl,-rpath,$ORIGIN/../mycode_amd",
           "-Wl,-rpath-link,/opt/rocm/lib", "-o", HOST_TEST_BIN + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    _run(cmd)
    os.replace(HOST_TEST_BIN + ".tmp", HOST_TEST_BIN)
    return HOST_TEST_BIN


def build(verbose: bool = False) -> str:
    """Compile every source and link the shared library (and the C++ host
    test program); returns the library's path."""
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src in sorted(os.listdir(CSRC)):
        path = os.path.join(CSRC, src)
        obj = os.path.join(BUILD, src + ".o")
        if src.endswith(".cpp"):
            cmd = ["g++", *COMMON, "-O2", "-c", path, "-o", obj, "-I/opt/rocm/include",
                   "-D__HIP_PLATFORM_AMD__"]
        elif src.endswith(".hip"):
            cmd = [HIPCC, *HIP_FLAGS, "-c", path, "-o", obj]
        else:
            continue
        if verbose:
            print(" ".join(cmd), flush=True)
        _run(cmd)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    if os.path.exists(HOST_TEST_SRC):
        build_host_test(verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
