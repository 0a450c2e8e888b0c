"""The built library's kernels, read from its own gfx950 code object (no GPU needed).

The planner's team pass is a cooperative launch: every resident work-group stays alive until
the batch ends, so its private memory (scratch) is needed for every lane of the GPU at once.
Two builds whose round loop the compiler had moved out of line, with a copy of the kernel
arguments in private memory (1456 and 1488 bytes per lane in the 16-dimension team kernel,
about 195 MB for the whole GPU), are the only two whose planner launches ever aborted.  These
tests keep every planner kernel's private segment small and check the metadata the library's
launch guard relies on.
"""

from __future__ import annotations

import re
import shutil
import subprocess
from pathlib import Path

import pytest

from prrax_amd import _lib

LLVM = Path("/opt/rocm/llvm/bin")
MAX_PLANNER_SCRATCH = 512  # bytes per lane; 512 x 512 threads x 256 work-groups = 64 MiB


def _kernel_metadata(tmp_path: Path) -> dict:
    bundler, readelf = LLVM / "clang-offload-bundler", LLVM / "llvm-readelf"
    if not (bundler.exists() and readelf.exists() and shutil.which("objcopy")):
        pytest.skip("ROCm LLVM tools not available")
    fat = tmp_path / "fatbin.bin"
    subprocess.run(["objcopy", f"--dump-section=.hip_fatbin={fat}", str(_lib.LIB_PATH), str(tmp_path / "strip.so")],
                   check=True)
    # one offload bundle per translation unit, concatenated in the section
    data = fat.read_bytes()
    magic = b"__CLANG_OFFLOAD_BUNDLE__"
    starts = [m.start() for m in re.finditer(re.escape(magic), data)]
    notes = ""
    for k, a in enumerate(starts):
        part = tmp_path / f"bundle{k}.bin"
        part.write_bytes(data[a:starts[k + 1] if k + 1 < len(starts) else len(data)])
        co = tmp_path / f"gfx950_{k}.co"
        subprocess.run([str(bundler), "--unbundle", "--type=o", f"--input={part}",
                        "--targets=hipv4-amdgcn-amd-amdhsa--gfx950", f"--output={co}"], check=True)
        notes += subprocess.run([str(readelf), "--notes", str(co)], check=True, capture_output=True, text=True).stdout
    out, cur = {}, None
    for line in notes.splitlines():
        m = re.match(r"\s+\.name:\s+(\S+)", line)
        if m:
            cur = m.group(1)
            out[cur] = {}
            continue
        m = re.match(r"\s+\.(private_segment_fixed_size|vgpr_count|vgpr_spill_count|sgpr_spill_count):\s+(\d+)", line)
        if m and cur:
            out[cur][m.group(1)] = int(m.group(2))
    return out


def test_planner_kernels_keep_private_memory_small(tmp_path):
    _lib.lib()  # the library next to the sources, built from them
    meta = _kernel_metadata(tmp_path)
    planner = {k: v for k, v in meta.items() if "plan_pass" in k}
    assert len(planner) == 4, sorted(meta)  # pass 1 and pass 2, for dim <= 8 and dim <= 16
    for name, m in planner.items():
        assert m["private_segment_fixed_size"] <= MAX_PLANNER_SCRATCH, (name, m)
        assert m["vgpr_count"] <= 256, (name, m)
    for name, m in meta.items():
        assert "private_segment_fixed_size" in m, name
