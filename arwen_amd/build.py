"""Build ``libarwen_dycore.so`` for gfx950 with hipcc, in-tree.

The library is compiled with ``-ffp-contract=off``: every expression
rounds exactly where its source expression says, which is what lets the
device step agree with the NumPy restatement of the reference step to the
bit.  Each kernel source is compiled twice, as the float32 build (entry
points ``arw_*``) and the FP64 build of the same discretisation
(``arw64_*``); the four objects link into one library.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
SOURCES = [PKG / "csrc" / "dycore.hip", PKG / "csrc" / "march.hip",
           PKG / "csrc" / "acoustic_adj.hip", PKG / "csrc" / "advect_adj.hip",
           PKG / "csrc" / "stage_adj.hip", PKG / "csrc" / "health.hip",
           PKG / "csrc" / "smag.hip"]
HEADERS = [PKG / "csrc" / "common.h", PKG / "csrc" / "march.h",
           ROOT / "include" / "arwen_dycore.h"]
OUT = PKG / "libarwen_dycore.so"

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{ROOT / 'include'}",
]
OBJ_DIR = PKG / "_build"


def hipcc() -> str:
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    cand = Path(rocm) / "bin" / "hipcc"
    return os.fspath(cand) if cand.exists() else "hipcc"


def build(verbose: bool = False) -> Path:
    """Compile every HIP source in both precisions and link the in-tree
    shared library."""
    OBJ_DIR.mkdir(exist_ok=True)
    jobs, objs = [], []
    for src in SOURCES:
        for f64 in (0, 1):
            obj = OBJ_DIR / f"{src.stem}{64 if f64 else 32}.o"
            cmd = [hipcc(), *HIPCC_FLAGS, f"-DARW_F64={f64}", "-c",
                   os.fspath(src), "-o", os.fspath(obj)]
            if verbose:
                print(" ".join(cmd), flush=True)
            jobs.append((cmd, subprocess.Popen(
                cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                text=True)))
            objs.append(obj)
    failed = False
    for cmd, proc in jobs:
        out, err = proc.communicate()
        if proc.returncode != 0:
            sys.stderr.write(out + err)
            failed = True
    if failed:
        raise RuntimeError("hipcc failed")
    tmp = OUT.with_suffix(".so.tmp")
    link = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o",
            os.fspath(tmp), *map(os.fspath, objs)]
    if verbose:
        print(" ".join(link), flush=True)
    res = subprocess.run(link, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc link failed ({res.returncode})")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
