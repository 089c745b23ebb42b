"""Host build of the device field code with bounds checks on every table, list and
fixed-point index (tests/field_check.cpp): generator, partial table and full table give
bit-identical fields; support lists keep the exact path's support decisions and hold
every impulse within reach of their sub-cell, in the reference's order; the
multi-resolution level index, scales and weights are the float formula's; the 1D
kernel's cells made by PCG jump-ahead equal the sequential generator's."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_field_tables_and_support_lists_on_the_host(tmp_path):
    exe = tmp_path / "field_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", "-Wno-unknown-pragmas", "-ffp-contract=off",
                    os.path.join(HERE, "field_check.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe), "2000", "101"], capture_output=True, text=True)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    lines = out.stdout.split("\n")
    assert lines[0] == "table_mismatch 0"
    assert "list_missing 0 list_extra 0" in out.stdout
    assert "level_mismatch 0 weight_mismatch 0" in out.stdout
    assert "cells1d_mismatch 0 of" in out.stdout
    assert "coop_mismatch 0 of" in out.stdout
    assert out.stdout.strip().endswith("OK")
