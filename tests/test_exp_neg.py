"""Accuracy of the table-driven exponentials (fml_amd/csrc/exp_neg.h), measured
on the host: the header is compiled with g++ and compared with long-double
expl -- exp_neg (the ARAD overlap's) on its whole domain (-16, 0] on a
2,000,000-point grid and on both sides of every argument-reduction boundary;
exp_neg_wide (the atom-pair kernels') on [-760, 1e-9], including the switch to
the library exp at -707, the subnormal range and tiny positive arguments.
The bar: at most one ulp (the table value's rounding plus the final fused
multiply-add's)."""
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
    lines = subprocess.run([exe, "2000000"], capture_output=True, text=True, check=True).stdout.splitlines()
    count, worst, worst_x = lines[0].split()
    assert int(count) > 2_000_000
    assert float(worst) <= 1.0, f"exp_neg: {worst} ulp at x = {worst_x}"
    count, worst, worst_x = lines[1].split()
    assert int(count) > 2_000_000
    assert float(worst) <= 1.0, f"exp_neg_wide: {worst} ulp at x = {worst_x}"
