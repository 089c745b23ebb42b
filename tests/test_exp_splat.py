"""The 1D kernel's splat exponential (exp_splat_d) rounds to the same float as the C
library's double exp on the splat argument range (host build of the device header)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_exp_splat_matches_libm(tmp_path):
    exe = tmp_path / "exp_splat_check"
    subprocess.run(["g++", "-O2", "-std=c++17", os.path.join(HERE, "exp_splat_check.cpp"), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe), "20000000"], check=True, capture_output=True, text=True).stdout.split()
    n, bad, maxrel = int(out[0]), int(out[1]), float(out[2])
    assert n == 20000000
    assert bad == 0
    assert maxrel <= 2.3e-16
