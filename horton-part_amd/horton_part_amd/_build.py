"""Compile ``libhortonpart.so`` for gfx950 with hipcc (in-tree)."""

import os
import shutil
import subprocess
import tempfile
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
    """Build the library; returns its path.  Raises on compiler errors.

    The translation units compile in parallel to objects in a temporary
    directory, then link into ``lib/libhortonpart.so``."""
    LIB.parent.mkdir(parents=True, exist_ok=True)
    tmp = LIB.with_suffix(".so.tmp")
    sources = [CSRC / "hortonpart.hip", CSRC / "hp_batch.hip"]
    compile_flags = [f for f in HIPCC_FLAGS if f != "-shared"]
    with tempfile.TemporaryDirectory(prefix="hortonpart-build-") as work:
        objs, procs = [], []
        for src in sources:
            obj = os.path.join(work, src.stem + ".o")
            cmd = [hipcc(), *compile_flags, f"-I{INCLUDE}", f"-I{CSRC}", "-c", "-o", obj, str(src)]
            if verbose:
                print(" ".join(cmd))
            procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
            objs.append(obj)
        errors = []
        for src, proc in procs:
            _, err = proc.communicate()
            if proc.returncode != 0:
                errors.append(f"{src.name}:\n{err[-4000:]}")
        if errors:
            raise RuntimeError("hipcc failed:\n" + "\n".join(errors))
        cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *objs]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError("link failed:\n" + res.stderr[-4000:])
    os.replace(tmp, LIB)
    return LIB
