"""Builds the native library libasfdtd.so in-tree with hipcc for gfx950.

Flags that fix the numerics: no contraction of a * b + c into fused multiply-adds
(-ffp-contract=off, as the reference builds, CMakeLists.txt:27-31), single-precision
denormals flushed to zero on the device (the reference runs with FTZ/DAZ,
fdtd.hpp:314-336), and no packed FP32 instructions.
"""

from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PACKAGE = Path(__file__).resolve().parent
ROOT = PACKAGE.parent
CSRC = PACKAGE / "csrc"
LIBRARY = PACKAGE / "libasfdtd.so"

HIP_SOURCES = ["fdtd_step.hip", "responses.hip", "setup_device.hip"]
HOST_SOURCES = ["setup.cpp", "fitted.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("AS_OFFLOAD_ARCH", "gfx950")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", f"-I{ROOT / 'include'}", f"-I{CSRC}"]
DEVICE = [f"--offload-arch={ARCH}", "-fgpu-flush-denormals-to-zero",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def build(verbose: bool = False) -> Path:
    """Compiles every source (always) and links libasfdtd.so; returns its path."""
    out_dir = PACKAGE / "_build"
    out_dir.mkdir(exist_ok=True)
    objects = []
    for name in HIP_SOURCES:
        src = CSRC / name
        obj = out_dir / (src.stem + ".o")
        cmd = [HIPCC, "-c", *COMMON, *DEVICE, "-x", "hip", str(src), "-o", str(obj)]
        _run(cmd, verbose)
        objects.append(obj)
    for name in HOST_SOURCES:
        src = CSRC / name
        if not src.exists():
            continue
        obj = out_dir / (src.stem + ".o")
        cmd = [HIPCC, "-c", *COMMON, "-Wall", "-Wextra", str(src), "-o", str(obj)]
        _run(cmd, verbose)
        objects.append(obj)
    tmp = LIBRARY.with_suffix(".so.tmp")
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", *map(str, objects), "-o", str(tmp), "-lpthread"]
    _run(cmd, verbose)
    os.replace(tmp, LIBRARY)
    return LIBRARY


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build failed: {' '.join(cmd)}")


if __name__ == "__main__":
    print(build(verbose=True))
