"""Accuracy of the ARAD overlap's exponential (fml_amd/csrc/exp_neg.h) on its
whole argument domain (-16, 0], measured on the host: the header is compiled
with g++ and compared with long-double expl on a 2,000,000-point grid and on
both sides of every argument-reduction boundary.  The bar: at most one ulp
(the table value's rounding plus the final fused multiply-add's)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_exp_neg_within_one_ulp(tmp_path):
    exe = str(tmp_path / "exp_neg_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "fml_amd", "csrc"),
                    os.path.join(ROOT, "tests", "exp_neg_check.cpp"), "-o", exe], check=True)
    out = subprocess.run([exe, "2000000"], capture_output=True, text=True, check=True).stdout.split()
    count, worst, worst_x = int(out[0]), float(out[1]), float(out[2])
    assert count > 2_000_000
    assert worst <= 1.0, f"{worst} ulp at x = {worst_x}"
