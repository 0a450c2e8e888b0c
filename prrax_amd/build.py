"""Builds the HIP planner library for gfx950.

``python -m prrax_amd.build`` (or ``__graft_entry__.build()``) compiles ``prrax_amd/csrc/*.hip``
with hipcc into ``prrax_amd/_prrtc_planner_lib.so`` -- the file name the reference loads
(cuda-rrtc/jax/prrtc.py:48).

Floating point: ``-ffp-contract=off`` (every FMA is an explicit ``__builtin_fmaf``), IEEE
division and square root, denormals kept, and packed FP32 instructions disabled.
"""

from __future__ import annotations

import hashlib
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "_prrtc_planner_lib.so"
ARCH = os.environ.get("PRRTC_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fPIC",
    "-shared",
    "-Wall",
]


def sources() -> list[Path]:
    return sorted(CSRC.glob("*.hip"))


def source_files() -> list[Path]:
    """Every file the library is compiled from: the HIP sources, their headers and the C ABI."""
    return sorted([*CSRC.glob("*.hip"), *CSRC.glob("*.h"), *(ROOT / "include").glob("*.h")])


def source_hash() -> str:
    """First 16 hex digits of a SHA-256 over the library's source files (names and contents);
    the build passes it to ``prrtc_version()``, so a loaded library can be matched to the sources
    next to it."""
    h = hashlib.sha256()
    for f in source_files():
        h.update(f.relative_to(ROOT).as_posix().encode())
        h.update(b"\0")
        h.update(f.read_bytes())
        h.update(b"\0")
    return h.hexdigest()[:16]


def build(verbose: bool = False, extra: list[str] | None = None) -> Path:
    cmd = ["hipcc", *HIP_FLAGS, f'-DPRRTC_SRC_HASH="{source_hash()}"', *(extra or []), "-I", str(ROOT / "include"),
           "-I", str(CSRC), "-o", str(LIB) + ".tmp", *map(str, sources())]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(str(LIB) + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True, extra=[a for a in sys.argv[1:] if a != "-v"])
