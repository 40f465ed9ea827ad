"""The FP64 exponentials of the device code (Tang's table method,
horton-part_amd/csrc/hp_exp_tab.h) against long double on the CPU: the same
source compiled with g++, swept over [-745, 0] (exp) and the same range in
units of ln2/64 (the weights pass's scaled form)."""

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
    # exp(x) and the scaled 2^(y/64): each within 1.5 ulp everywhere, within
    # 1 ulp for > 99.5% of arguments
    for prefix in ("", "scaled: "):
        m = re.search(r"^" + prefix + r"n=\d+ max_ulp=([0-9.]+) .* over_1ulp=(\d+)", out, re.M)
        assert m, out
        max_ulp, over1 = float(m.group(1)), int(m.group(2))
        assert max_ulp < 1.5, out
        assert over1 < 0.005 * 5_000_000, out
    assert "exp(0)=1 exp(-1e6)=0" in out
    assert "scaled: 2^0=1 2^(-1e9/64)=0" in out
