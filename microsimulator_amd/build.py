"""Builds the HIP backend library in-tree: microsimulator_amd/libmsim_hip.so.

hipcc compiles the kernels for gfx950 only. Flags:
  -ffp-contract=off      every float/double operation rounds on its own, as in
                         the reference's CPU build, so FP32 assembly and the
                         FP32 transport step match it bit for bit;
  -Xclang -target-feature -Xclang -packed-fp32-ops
                         no packed FP32 instructions.
Run as ``python -m microsimulator_amd.build`` (or through __graft_entry__.build).
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libmsim_hip.so")
BUILD = os.path.join(HERE, "_build")

SOURCES = ["flow.hip", "mg.hip", "adr.hip", "adr3.hip", "learned.hip", "batch.hip", "ports.hip", "implicit.hip",
           "coupled.hip", "colony.hip", "nutrient.hip", "stream.hip", "resolved.hip", "capi.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MSIM_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3", "-std=c++20", "-fPIC", "-ffp-contract=off",
    f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}",
]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *COMMON, *DEVICE, "-c", os.path.join(CSRC, src), "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC, *COMMON, "-x", "hip", *DEVICE, "-c", os.path.join(CSRC, src), "-o", obj]
    subprocess.run(cmd, check=True)
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as pool:
        objects = list(pool.map(_compile, SOURCES))
    tmp = OUT + ".tmp"
    subprocess.run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objects],
                   check=True)
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT}")
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
