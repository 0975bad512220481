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
HEADERS = ["ql_fft.hpp", "ql_kernels.hpp", "ql_mode.hpp", "ql_adjoint.hpp", "ql_noise.hpp", "ql_runtime.hpp", "host/ql_model.hpp"]
HOST_TEST_SRC = os.path.join(ROOT, "tests", "cpp", "test_host_adapter.cpp")
HOST_TEST_BIN = os.path.join(ROOT, "tests", "cpp", "test_host_adapter")


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall", "-Wno-unused-result"]


def build(verbose: bool = False) -> str:
    """Compiles the sources to objects in parallel (one hipcc each), then links
    libql_dns.so."""
    import tempfile
    tmp = LIB + ".tmp"
    with tempfile.TemporaryDirectory(prefix="ql_build_") as odir:
        procs = []
        for src in SOURCES:
            obj = os.path.join(odir, os.path.basename(src) + ".o")
            cmd = [hipcc()] + FLAGS + ["-I", os.path.join(ROOT, "include"), "-c", "-o", obj,
                                       os.path.join(CSRC, src)]
            if verbose:
                print(" ".join(cmd), flush=True)
            procs.append((obj, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        failed = False
        for obj, pr in procs:
            out, _ = pr.communicate()
            if pr.returncode != 0:
                sys.stderr.write(out)
                failed = True
        if failed:
            raise RuntimeError("hipcc failed compiling libql_dns.so")
        cmd = [hipcc(), "--offload-arch=gfx950", "-fPIC", "-shared", "-o", tmp] + [o for o, _ in procs]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("hipcc failed linking libql_dns.so")
    os.replace(tmp, LIB)
    io = build_io(verbose)
    build_host_test(verbose)
    if io:
        build_io_test(verbose)
    return LIB


IO_TEST_SRC = os.path.join(ROOT, "tests", "cpp", "test_io_restart.cpp")
IO_TEST_BIN = os.path.join(ROOT, "tests", "cpp", "test_io_restart")


def build_io_test(verbose: bool = False) -> str:
    """g++ build of the C++ dump / restart driver over FullSave_Load."""
    cmd = ["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), IO_TEST_SRC, "-L", HERE, "-lql_io",
           "-lql_dns", "-Wl,-rpath,$ORIGIN/../../ql_dns_amd", "-o", IO_TEST_BIN]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("g++ failed building tests/cpp/test_io_restart")
    return IO_TEST_BIN


IO_LIB = os.path.join(HERE, "libql_io.so")
IO_SRC = os.path.join(CSRC, "io", "ql_io.cpp")
HDF5_PREFIXES = ("/opt/conda", "/usr", "/usr/local")


def hdf5_prefix():
    for pre in HDF5_PREFIXES:
        if os.path.exists(os.path.join(pre, "include", "hdf5.h")):
            return pre
    return None


def build_io(verbose: bool = False):
    """g++ build of libql_io.so (HDF5 dump/restart and the .dat time series),
    linked against libql_dns.so and the HDF5 C library.  Skipped (with a
    message) when no HDF5 installation is found."""
    pre = hdf5_prefix()
    if pre is None:
        sys.stderr.write("ql_dns_amd.build: HDF5 not found; libql_io.so not built\n")
        return None
    # only the HDF5 library itself is reached through a private link
    # directory; putting the HDF5 prefix's lib/ on the run path would also
    # pull its (older) C++ runtime into the process
    libdir = os.path.join(pre, "lib")
    real = os.path.join(libdir, "libhdf5.so")
    soname = next((f for f in sorted(os.listdir(libdir)) if f.startswith("libhdf5.so.")), None)
    linkdir = os.path.join(HERE, "_hdf5")
    os.makedirs(linkdir, exist_ok=True)
    if soname:
        dst = os.path.join(linkdir, soname)
        if os.path.lexists(dst):
            os.remove(dst)
        os.symlink(os.path.join(libdir, soname), dst)
    tmp = IO_LIB + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(pre, "include"), IO_SRC, "-L", HERE, "-lql_dns", real,
           "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,$ORIGIN/_hdf5", "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("g++ failed building libql_io.so")
    os.replace(tmp, IO_LIB)
    return IO_LIB


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
