"""Build ``libarwen_dycore.so`` for gfx950 with hipcc, in-tree.

The library is compiled with ``-ffp-contract=off``: every float32
expression rounds exactly where its source expression says, which is what
lets the device step agree with the float32 restatement of the reference
step to the bit.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
SOURCES = [PKG / "csrc" / "dycore.hip", PKG / "csrc" / "march.hip"]
HEADERS = [PKG / "csrc" / "common.h", PKG / "csrc" / "march.h",
           ROOT / "include" / "arwen_dycore.h"]
OUT = PKG / "libarwen_dycore.so"

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{ROOT / 'include'}",
]


def hipcc() -> str:
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    cand = Path(rocm) / "bin" / "hipcc"
    return os.fspath(cand) if cand.exists() else "hipcc"


def build(verbose: bool = False) -> Path:
    """Compile every HIP source into the in-tree shared library."""
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [hipcc(), *HIPCC_FLAGS, "-o", os.fspath(tmp),
           *map(os.fspath, SOURCES)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed ({res.returncode})")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
