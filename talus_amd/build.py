"""Builds libtalus_mi.so (gfx950 HIP kernels + host C++ + C ABI) in-tree.

Usage: python -m talus_amd.build   (or talus_amd.build.build())
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libtalus_mi.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
SOURCES = ["kernels.hip", "factor.hip", "context.hip", "lattice.hip", "screen.hip", "field.hip", "contact.hip", "mesh.cpp",
           "factor_symbolic.cpp", "tear.cpp", "surface.cpp", "hex_build.cpp"]
COMMON = [
    "-O3", "-fPIC", "-std=c++20", "-Wall", "-Wno-unused-function",
    f"-I{os.path.join(ROOT, 'include')}",
]
# Packed FP32 instructions are kept out of every kernel (the compiler would
# otherwise form v_pk_* FP32 ops from float code).
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-ffp-contract=fast"]
# contact generation rounds every float operation as the reference does:
# no contraction, correctly rounded division and square root
EXTRA = {"contact.hip": ["-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt"]}


def _compile(src: str) -> str:
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, src + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC, "-x", "hip", *DEVICE, *EXTRA.get(src, []), *COMMON, "-c", path, "-o", obj]
    else:
        # host topology and symbolic code: no contraction, so results match
        # the oracle's -ffp-contract=off build operation for operation
        cmd = [HIPCC, *COMMON, "-ffp-contract=off", "-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    warnings = [l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip()]
    if warnings:
        sys.stderr.write("\n".join(warnings) + "\n")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
