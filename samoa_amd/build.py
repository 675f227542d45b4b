"""Builds the native libraries of the package in-tree.

    python samoa_amd/build.py

  samoa_amd/lib/libswe_mi355x.so  HIP kernels + C ABI (hipcc, gfx950)
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")

HIP_SOURCES = ["swe_step.hip", "swe_api.hip"]
HIPCC_FLAGS = [
    "-O3", "--offload-arch=gfx950", "-fPIC", "-shared", "-std=c++17",
    # every operation rounds as written: bit-identical to the reference
    "-ffp-contract=off",
    # no packed FP32 instructions (some of their forms are unreliable on this hardware)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def build_hip(verbose: bool = False) -> str:
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libswe_mi355x.so")
    cmd = ["hipcc", *HIPCC_FLAGS, f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}",
           *[os.path.join(CSRC, s) for s in HIP_SOURCES], "-o", out + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed")
    if verbose:
        sys.stderr.write("\n".join(l for l in r.stderr.splitlines() if "packed-fp32" not in l))
    os.replace(out + ".tmp", out)
    return out


def build_all(verbose: bool = False) -> None:
    build_hip(verbose)


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
    print("built", os.path.join(LIB, "libswe_mi355x.so"))
