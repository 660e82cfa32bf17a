"""Build libhydra_amd.so (HIP, gfx950) in-tree.

    python -m hydra_amd.build

Every HIP source is compiled for gfx950 with:
  -ffp-contract=off       no FMA contraction: the plain FP64 path must round
                          exactly like the reference's x86-64 build
  -packed-fp32-ops off    packed FP32 VALU ops are kept out of the code object
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libhydra_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("HYDRA_AMD_ARCH", "gfx950")

SOURCES = ["tables.cpp", "fft.hip", "con2grid.hip", "seqsum.hip", "nodes.hip", "step.hip",
           "renode.hip", "congen.hip", "surgery.hip", "capi.hip", "cnn.hip", "diag.hip"]

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value",
]


def build(verbose: bool = False) -> str:
    objs = []
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(objdir, src + ".o")
        lang = [] if src.endswith(".cpp") else ["-x", "hip"]
        cmd = [HIPCC, *lang, *FLAGS, "-c", path, "-o", obj]
        if src.endswith(".cpp"):
            cmd = [HIPCC, "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
                   "-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
