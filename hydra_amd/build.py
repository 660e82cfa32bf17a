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


def _compile(src: str, objdir: str, verbose: bool) -> str:
    path = os.path.join(CSRC, src)
    obj = os.path.join(objdir, src + ".o")
    if src.endswith(".cpp"):
        cmd = [HIPCC, "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
               "-c", path, "-o", obj]
    else:
        cmd = [HIPCC, "-x", "hip", *FLAGS, "-c", path, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return obj


def build(verbose: bool = False, jobs: int | None = None) -> str:
    """Compile every source (in parallel, at most 8 at a time) and link."""
    from concurrent.futures import ThreadPoolExecutor
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 1)
    # longest compiles first so the pool finishes together
    order = sorted(SOURCES, key=lambda f: -os.path.getsize(os.path.join(CSRC, f)))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        done = dict(zip(order, ex.map(lambda f: _compile(f, objdir, verbose), order)))
    objs = [done[f] for f in SOURCES]
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
