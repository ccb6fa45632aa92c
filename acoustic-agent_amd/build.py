"""Compile the in-tree HIP library ``libaa_tracer.so`` for gfx950.

The library holds the CDNA4 kernels and the C ABI declared in
``include/aa_tracer.h``.  It is built in place (next to this file) so that it
travels with the source tree; nothing is installed.
"""
from __future__ import annotations

import os
import shutil
import subprocess
from pathlib import Path

PACKAGE_DIR = Path(__file__).resolve().parent
REPO_DIR = PACKAGE_DIR.parent
CSRC = PACKAGE_DIR / "csrc"
LIB_PATH = PACKAGE_DIR / "libaa_tracer.so"
SOURCES = ("aa_tracer.hip", "aa_bvh.cpp", "aa_batch.hip", "aa_fdn.cpp")

# Device code: IEEE division/sqrt, no contraction (the binary32 trajectory
# must round like the CPU restatement), no packed-FP32 instructions.
HIP_FLAGS = (
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
)


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(path).exists():
        raise RuntimeError("hipcc not found; the ROCm toolchain is required to build libaa_tracer.so")
    return path


def build(verbose: bool = False) -> Path:
    """Compile every source and link ``libaa_tracer.so``; returns its path."""
    build_dir = PACKAGE_DIR / "_build"
    build_dir.mkdir(exist_ok=True)
    objects = []
    for name in SOURCES:
        src = CSRC / name
        if not src.exists():
            continue
        obj = build_dir / (name + ".o")
        lang = ["-x", "hip"] if name.endswith(".hip") else []
        cmd = [hipcc(), *HIP_FLAGS, *lang, "-I", str(REPO_DIR / "include"), "-c", str(src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objects.append(str(obj))
    tmp = LIB_PATH.with_suffix(".so.tmp")
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *objects, "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))
