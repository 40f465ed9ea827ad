"""Compile ``libhortonpart.so`` for gfx950 with hipcc (in-tree)."""

import os
import shutil
import subprocess
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent.parent
CSRC = PKG.parent / "csrc"
INCLUDE = ROOT / "include"
LIB = PKG / "lib" / "libhortonpart.so"

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # no packed-FP32 instructions in generated code
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
]


def hipcc():
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(exe).exists():
        raise RuntimeError("hipcc not found")
    return exe


def build(verbose=False):
    """Build the library; returns its path.  Raises on compiler errors."""
    LIB.parent.mkdir(parents=True, exist_ok=True)
    tmp = LIB.with_suffix(".so.tmp")
    sources = [str(CSRC / "hortonpart.hip"), str(CSRC / "hp_batch.hip")]
    cmd = [hipcc(), *HIPCC_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", "-o", str(tmp), *sources]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + res.stderr[-4000:])
    if verbose:
        print(" ".join(cmd))
    os.replace(tmp, LIB)
    return LIB
