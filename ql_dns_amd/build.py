"""Builds libql_dns.so (HIP, gfx950) in-tree.

    python -m ql_dns_amd.build          # or ql_dns_amd.build.build()

The library is compiled with hipcc for gfx950 only.  Packed FP32 instructions
are disabled (``-target-feature -packed-fp32-ops``) for all device code.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libql_dns.so")
SOURCES = ["ql_dns.hip", "ql_adjoint.hip", "host/ql_model.cpp"]
HEADERS = ["ql_fft.hpp", "ql_kernels.hpp", "ql_mode.hpp", "ql_adjoint.hpp", "ql_runtime.hpp", "host/ql_model.hpp"]
HOST_TEST_SRC = os.path.join(ROOT, "tests", "cpp", "test_host_adapter.cpp")
HOST_TEST_BIN = os.path.join(ROOT, "tests", "cpp", "test_host_adapter")


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def build(verbose: bool = False) -> str:
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Wall", "-Wno-unused-result",
           "-I", os.path.join(ROOT, "include"),
           "-o", tmp] + [os.path.join(CSRC, s) for s in SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libql_dns.so")
    os.replace(tmp, LIB)
    build_host_test(verbose)
    return LIB


def build_host_test(verbose: bool = False) -> str:
    """g++ build of the C++ driver that exercises the Model/Integrator mirror."""
    cmd = ["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), HOST_TEST_SRC,
           "-L", HERE, "-lql_dns", "-Wl,-rpath,$ORIGIN/../../ql_dns_amd", "-o", HOST_TEST_BIN]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("g++ failed building tests/cpp/test_host_adapter")
    return HOST_TEST_BIN


if __name__ == "__main__":
    print(build(verbose=True))
