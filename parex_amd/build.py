"""Build recipe for the in-tree shared libraries (run on the CPU host; hipcc
cross-compiles gfx950 code without a GPU).

  parex_amd/libpxgen.so   host graph ingest (gcc, OpenMP)
  parex_amd/libpxd.so     HIP engine + C ABI (hipcc --offload-arch=gfx950)
  parex_amd/libpxh.so     expander hierarchy + normalized cut (g++, OpenMP, links libpxd)
  parex_amd/bin/parex     CLI drop-in (g++, links both)

The oracle and the C++ test programs (test infrastructure) are built by
oracle/Makefile and tests/cpp/Makefile.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INC = os.path.join(ROOT, "include")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # IEEE results the oracle reproduces: no contraction into fused multiply-add
    "-ffp-contract=off",
    # packed FP32 instructions are kept out of all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if out.strip():
        print(out)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}")


def build_pxgen():
    _run(["gcc", "-O2", "-fPIC", "-fopenmp", "-std=gnu11", "-Wall", f"-I{INC}", "-shared",
          "-o", os.path.join(PKG, "libpxgen.so"), os.path.join(CSRC, "pxgen.c"), "-lm"])


def build_engine():
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    _run([hipcc, *HIP_FLAGS, f"-I{INC}", f"-I{CSRC}", os.path.join(CSRC, "pxd_engine.hip"),
          "-o", os.path.join(PKG, "libpxd.so")])


def build_cli():
    """libparex_cuda.so: the CudaDeviceManager facade; bin/parex: the CLI."""
    _run(["g++", "-O2", "-std=c++20", "-Wall", "-fPIC", "-shared", f"-I{INC}", os.path.join(CSRC, "facade.cpp"),
          "-o", os.path.join(PKG, "libparex_cuda.so"), f"-L{PKG}", "-lpxd", "-Wl,-rpath,$ORIGIN"])
    os.makedirs(os.path.join(PKG, "bin"), exist_ok=True)
    _run(["g++", "-O2", "-std=c++20", "-Wall", f"-I{INC}", os.path.join(CSRC, "cli.cpp"),
          "-o", os.path.join(PKG, "bin", "parex"), f"-L{PKG}", "-lparex_cuda", "-lpxd", "-lpxgen",
          "-Wl,-rpath,$ORIGIN/.."])


def build_hierarchy():
    """libpxh.so: the hierarchy and normalized cut over the engine (pxh.h)."""
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-fPIC", "-shared", "-ffp-contract=off", "-fopenmp", f"-I{INC}",
          os.path.join(CSRC, "pxh.cpp"), "-o", os.path.join(PKG, "libpxh.so"), f"-L{PKG}", "-lpxd",
          "-Wl,-rpath,$ORIGIN"])


def build_oracle():
    _run(["make", "-C", os.path.join(ROOT, "oracle"), "-B", "all"])


def build_test_programs():
    """tests/cpp: the CudaDeviceManager facade stepped against the oracle."""
    _run(["make", "-C", os.path.join(ROOT, "tests", "cpp"), "-B", "all"])


def build(oracle: bool = True):
    build_pxgen()
    build_engine()
    build_cli()
    build_hierarchy()
    if oracle:
        build_oracle()
        build_test_programs()


if __name__ == "__main__":
    build(oracle="--no-oracle" not in sys.argv)
