"""Compiles the HIP library ``hicar_amd/libicar_adv.so`` for gfx950.

The library is built in-tree with ``hipcc``; object files go to
``hicar_amd/_build``.  All kernels are compiled without floating-point
contraction so that they keep the reference's IEEE operation order (FP32
division stays correctly rounded, HIP's default).  Packed FP32 instructions
are disabled for all device code.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libicar_adv.so")
ARCH = os.environ.get("HICAR_AMD_ARCH", "gfx950")

COMMON = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
# per translation unit: (source, extra flags)
UNITS = [
    ("kernels_exact.hip", ["-ffp-contract=off"]),
    ("kernels_fast.hip", ["-ffp-contract=off"]),
    ("kernels_dyn.hip", ["-ffp-contract=off"]),
    ("kernels_std.hip", ["-ffp-contract=off"]),
    ("api.hip", ["-ffp-contract=off"]),
    ("noalloc.hip", ["-ffp-contract=off"]),
    ("divcheck.hip", ["-ffp-contract=off"]),
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(
        line for line in proc.stdout.splitlines()
        # the host pass does not know the device-only target feature
        if "packed-fp32-ops" not in line and "is not a recognized feature" not in line
    )
    if proc.returncode != 0:
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip():
        sys.stderr.write(out + "\n")


def build(verbose: bool = False) -> str:
    """Compiles every translation unit and links the shared library."""
    cc = hipcc()
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src, extra in UNITS:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        cmd = [cc, *COMMON, *extra, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
