"""Builds the in-tree HIP library ``bajo_amd/libbajo_hip.so`` for gfx950.

Usage: ``python -m bajo_amd.build`` (or ``__graft_entry__.build()``).  Each
translation unit is compiled with hipcc in parallel, then linked into one
shared library.  FP contraction is off so every fused multiply-add in the
kernels is an explicit ``fmaf`` matching bajo's arithmetic; packed-FP32
instructions are disabled.
"""

from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libbajo_hip.so"
ARCH = os.environ.get("BAJO_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

COMMON_FLAGS = [
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-ffp-contract=off",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-result",
]
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
]

SOURCES = ["context.hip", "sort.hip", "lbvh.hip", "hploc.hip", "wide.hip", "trace.hip", "tlas.hip", "render.hip", "guide.hip",
           "scenes.cpp"]


def _compile(src: Path) -> Path:
    out = OBJ / (src.name + ".o")
    cmd = [HIPCC, *COMMON_FLAGS]
    if src.suffix == ".hip":
        cmd += ["-x", "hip", *HIP_FLAGS]
    cmd += ["-c", str(src), "-o", str(out)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{proc.stdout}\n{proc.stderr}")
    return out


def build(verbose: bool = False) -> Path:
    OBJ.mkdir(exist_ok=True)
    srcs = [CSRC / s for s in SOURCES]
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{proc.stdout}\n{proc.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
