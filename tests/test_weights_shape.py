"""Workgroup shapes and shell groups of the weights pass (hp_host.h): the
invariants the kernel relies on, checked by a host-only program built with
hipcc (tests/native/weights_shape_check.cpp)."""

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="needs hipcc")
def test_weights_shapes_and_groups(tmp_path):
    exe = tmp_path / "weights_shape_check"
    src = os.path.join(ROOT, "tests", "native", "weights_shape_check.cpp")
    inc = [f"-I{os.path.join(ROOT, 'include')}", f"-I{os.path.join(ROOT, 'horton-part_amd', 'csrc')}"]
    subprocess.run([HIPCC, "-x", "hip", "--offload-host-only", "-std=c++17", *inc, "-o", str(exe), src],
                   check=True, capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    assert out.strip() == "ok", out
