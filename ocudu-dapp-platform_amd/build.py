#!/usr/bin/env python3
"""Builds the HIP LDPC receive-chain library for gfx950 in-tree.

    python ocudu-dapp-platform_amd/build.py

Output: ocudu-dapp-platform_amd/lib/libocudu_hip_ldpc.so (C ABI of include/ocudu_hip_ldpc.h).
"""
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "lib"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-std=c++17", "-fPIC", f"-I{ROOT / 'include'}", f"-I{CSRC}", "-Wall",
          "-Wno-unused-function"]
# Packed FP32 VALU forms are disabled for every device compile.
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def run(cmd):
    print(" ".join(str(c) for c in cmd), flush=True)
    subprocess.run([str(c) for c in cmd], check=True)


def build(verbose=True):
    LIB.mkdir(exist_ok=True)
    objs = []
    obj_dir = LIB / "obj"
    obj_dir.mkdir(exist_ok=True)
    for src in sorted(CSRC.glob("*.hip")):
        o = obj_dir / (src.stem + ".o")
        run([HIPCC, *COMMON, *DEVICE, "-x", "hip", "-c", src, "-o", o])
        objs.append(o)
    for src in sorted(CSRC.glob("*.cpp")):
        o = obj_dir / (src.stem + ".o")
        run([HIPCC, *COMMON, "-c", src, "-o", o])
        objs.append(o)
    out = LIB / "libocudu_hip_ldpc.so"
    run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", out, *objs])
    # C++ adapters implementing the reference's class interfaces over the C ABI.
    adapters = LIB / "libocudu_hip_adapters.so"
    run([HIPCC, *COMMON, f"-I{PKG / 'host'}", "-shared", PKG / "host" / "hip_ldpc_adapters.cpp", "-o", adapters,
         f"-L{LIB}", "-locudu_hip_ldpc", f"-Wl,-rpath,{LIB}"])
    return out


def build_tests(oracle_lib):
    """The C++ tests (tests/cpp): the adapter chain test, linking the adapters and the CPU
    oracle, and the demultiplexer's host-side check."""
    out = ROOT / "tests" / "cpp" / "adapter_chain_test"
    run([HIPCC, *COMMON, f"-I{PKG / 'host'}", ROOT / "tests" / "cpp" / "adapter_chain_test.cpp", "-o", out,
         f"-L{LIB}", "-locudu_hip_adapters", "-locudu_hip_ldpc", f"-Wl,-rpath,{LIB}",
         f"-L{oracle_lib.parent}", "-loracle", f"-Wl,-rpath,{oracle_lib.parent}"])
    # Host-side replay of the device demultiplexer's arithmetic against the oracle.
    demux = ROOT / "tests" / "cpp" / "demux_rank_test"
    run([HIPCC, *COMMON, ROOT / "tests" / "cpp" / "demux_rank_test.cpp", "-o", demux, f"-L{LIB}",
         "-locudu_hip_ldpc", f"-Wl,-rpath,{LIB}", f"-L{oracle_lib.parent}", "-loracle",
         f"-Wl,-rpath,{oracle_lib.parent}"])
    return out


if __name__ == "__main__":
    build()
    sys.exit(0)
