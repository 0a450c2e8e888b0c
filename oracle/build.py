"""Build recipe for the CPU oracle (test infrastructure only).

Compiles ``oracle/prrtc_oracle.c`` into ``oracle/liboracle.so`` with ``-ffp-contract=off`` so
that every fused multiply-add is the explicit ``fmaf`` the source writes.

The reference's own planner cannot be built here: ``cuda-rrtc/CUDA/prrtc_planner.cu`` needs
nvcc, the CUDA runtime headers, jaxlib's ``xla/ffi/api/ffi.h`` and pyroffi's
``_collision_cuda_helpers.cuh`` (``cuda-rrtc/build.sh:34-65``), none of which exist in this
image, so there is no ``oracle/_ref`` build.
"""

from __future__ import annotations

import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRC = HERE / "prrtc_oracle.c"
OUT = HERE / "liboracle.so"
FLAGS = ["-O2", "-mavx2", "-mfma", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared", "-Wall"]


def build(verbose: bool = False) -> Path:
    cmd = ["gcc", *FLAGS, "-o", str(OUT), str(SRC), "-lm"]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    return OUT


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
