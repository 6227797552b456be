"""Build libmhs_amd.so in-tree for gfx950 (hipcc; no cmake, no JIT cache).

    python -m mhs_amd.build
"""
from __future__ import annotations

import os
import pathlib
import subprocess
import sys

HERE = pathlib.Path(__file__).resolve().parent
SRC = HERE / "csrc" / "kernels.hip"
SRC_WELD = HERE / "csrc" / "weld.hip"
OUT = HERE / "libmhs_amd.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
         # packed FP32 instructions are kept out of every HIP build
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         # fully unroll every `#pragma unroll` loop of known trip count: at the
         # default size limit the image edge integrals' 66-step series stays a
         # loop that loads each binomial coefficient and a dynamically indexed
         # accumulator from memory and waits on every step
         "-mllvm", "-pragma-unroll-threshold=65536",
         "-I", str(HERE.parent / "include")]


def build(verbose: bool = True) -> pathlib.Path:
    """Regenerate the image-family code if the table is newer, then compile."""
    from .gen import emit
    inc = HERE / "csrc" / "gen_image.inc"
    table = HERE / "data" / "mindlin_table.json"
    if not inc.exists() or inc.stat().st_mtime < table.stat().st_mtime:
        emit.main()
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [HIPCC, *FLAGS, "-o", str(tmp), str(SRC), str(SRC_WELD)]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed ({r.returncode})")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build())
