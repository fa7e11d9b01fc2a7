"""Build recipe for libspgemm_amd.so (gfx950) and its C++ test programs.

    python spgemm-hpc-thesis_amd/build.py   # library + the C++ shim test program

Device code is compiled with -ffp-contract=off: the engine's bit-exactness
rests on every product and every sum being rounded separately.  Packed FP32
instructions are disabled for all device code.
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIBDIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIBDIR, "libspgemm_amd.so")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SPGEMM_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
    f"-I{INCLUDE}", f"-I{CSRC}",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-pthread", "-ffp-contract=off", "-Wall", "-Wextra", f"-I{INCLUDE}"]

HIP_SOURCES = ["analyze.hip", "symbolic.hip", "numeric.hip", "fused.hip", "scan.hip", "engine.hip", "batch.hip", "accumulate.hip",
               "halo.hip", "noalloc.hip", "large_part.hip"]
CXX_SOURCES = ["host_matrix.cpp", "server.cpp", "validate.cpp"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ...")
    return r


def build_library(verbose: bool = False) -> str:
    """Compile every HIP/C++ source and link libspgemm_amd.so (always rebuilds)."""
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj], verbose)
        objs.append(obj)
    for src in CXX_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        _run(["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj], verbose)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", "-o", tmp, *objs], verbose)
    os.replace(tmp, LIB)
    return LIB


def build_shim_test(verbose: bool = False) -> str:
    """Compile the C++17 shim test program against the library."""
    src = os.path.join(PKG, "host", "shim_test.cpp")
    exe = os.path.join(LIBDIR, "shim_test")
    _run(["g++", "-O2", "-std=c++17", f"-I{INCLUDE}", f"-I{os.path.join(PKG, 'host')}", src, "-o", exe,
          f"-L{LIBDIR}", "-lspgemm_amd", "-Wl,-rpath,$ORIGIN"], verbose)
    return exe


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    if not shutil.which(HIPCC) and not os.path.exists(HIPCC):
        sys.exit("hipcc not found")
    print(build_library(a.verbose))
    print(build_shim_test(a.verbose))


if __name__ == "__main__":
    main()
