"""Builds disco2_amd/libdisco2.so (HIP kernels for gfx950 + host C ABI).

    python -m disco2_amd.build

Every object is compiled with -ffp-contract=off, so neither the host code
nor the device kernels contract a*b+c into an FMA: results then follow the
reference's IEEE operation order.  The initial-data plugins and the analytic
frame rotation are compiled by g++, the reference's compiler family, so that
the C library calls they make (pow, exp, ...) are folded or called as in the
reference build.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libdisco2.so")
ARCH = os.environ.get("DISCO_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wno-unused-variable",
          "-Wno-unused-function", "-Wno-unused-result"]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]
CXX = os.environ.get("DISCO_CXX", "g++")
SOURCES = ["kernels.hip", "unit.hip", "adjoint.hip", "diagnostics.hip", "checkpoint.hip", "engine.cpp", "adjoint_host.cpp",
           "diagnostics.cpp", "checkpoint.cpp", "h5.cpp", "capi.cpp", "sim_host.cpp", "initial_data.cpp", "frame.cpp"]
GCC_SOURCES = {"initial_data.cpp", "frame.cpp"}


def _compile(src: str) -> str:
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, src + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC, "-x", "hip", *DEVICE, *COMMON, "-c", path, "-o", obj]
    elif src in GCC_SOURCES:
        cmd = [CXX, "-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-c", path, "-o", obj]
    else:
        cmd = [HIPCC, *COMMON, "-D__HIP_PLATFORM_AMD__", "-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=4) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
