"""Build of the in-tree HIP library ``pyhype_amd/lib/libpyhype_hip.so`` for gfx950.

The library is compiled with hipcc directly (no JIT cache) so that the built ``.so`` travels with
the source tree.  Floating-point flags: ``-ffp-contract=off`` keeps every multiply and add a
separately rounded IEEE operation, which the bit-for-bit parity with the reference arithmetic
relies on; packed-FP32 instruction selection is disabled.
"""
from __future__ import annotations

import os
import subprocess
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "lib"
LIB = LIBDIR / "libpyhype_hip.so"
INCLUDE = PKG.parent / "include"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
]

SOURCES = ["phx_kernels.hip", "phx_adjoint.hip"]


def build_library(verbose: bool = False, extra_flags=None) -> Path:
    """Compile every HIP source into ``lib/libpyhype_hip.so`` (always recompiles)."""
    LIBDIR.mkdir(exist_ok=True)
    srcs = [str(CSRC / s) for s in SOURCES if (CSRC / s).exists()]
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, *HIP_FLAGS, *(extra_flags or []), "-shared", f"-I{INCLUDE}", *srcs, "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{res.stdout}\n{res.stderr}")
    if verbose and res.stderr:
        print(res.stderr)
    os.replace(tmp, LIB)
    return LIB


def build_test_library(verbose: bool = False) -> Path:
    """Compile the GPU tests' device numerics library (tests/native/libdevice_numerics.so)."""
    src = PKG.parent / "tests" / "native" / "device_numerics.hip"
    out = src.with_name("libdevice_numerics.so")
    tmp = out.with_suffix(".so.tmp")
    cmd = [HIPCC, *HIP_FLAGS, "-shared", f"-I{INCLUDE}", str(src), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build_library(verbose=True))
