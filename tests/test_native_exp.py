"""The FP64 exponential of the weights pass (Tang's table method,
horton-part_amd/csrc/hp_exp_tab.h) against long-double exp on the CPU: the
same source compiled with g++, swept over [-745, 0]."""

import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_exp_table_accuracy(tmp_path):
    exe = tmp_path / "exp_tab_check"
    src = os.path.join(ROOT, "tests", "native", "exp_tab_check.cpp")
    inc = os.path.join(ROOT, "horton-part_amd", "csrc")
    subprocess.run(["g++", "-O2", "-std=c++17", f"-I{inc}", "-o", str(exe), src], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    m = re.search(r"max_ulp=([0-9.]+) .* over_1ulp=(\d+)", out)
    assert m, out
    max_ulp, over1 = float(m.group(1)), int(m.group(2))
    # correctly rounded within 1.5 ulp everywhere, within 1 ulp for > 99.5% of arguments
    assert max_ulp < 1.5, out
    assert over1 < 0.005 * 5_000_000, out
    assert "exp(0)=1 exp(-1e6)=0" in out
