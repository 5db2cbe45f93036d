"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only):

* the C oracle (oracle/fml_oracle.c), built with -fsanitize=address,undefined
  -fno-sanitize-recover=all (make -C oracle sanitize), runs the whole oracle
  test file in a child interpreter with the sanitizer runtimes preloaded;
* the exp/sin header shared with the device code (fml_amd/csrc/exp_neg.h),
  compiled with the same flags into its host check program.

Any sanitizer report aborts the child, which fails the test."""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"]


def _runtime(name):
    if shutil.which("gcc") is None:
        return None
    p = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
    return p if os.path.isabs(p) and os.path.exists(p) else None


def _env(extra=None):
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=1"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    env.update(extra or {})
    return env


@pytest.mark.skipif(_runtime("libasan.so") is None or _runtime("libubsan.so") is None,
                    reason="needs gcc's ASan and UBSan runtimes")
def test_oracle_under_sanitizers():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True)
    lib = os.path.join(ROOT, "oracle", "_san", "liboracle_fml_san.so")
    env = _env({"LD_PRELOAD": _runtime("libasan.so") + ":" + _runtime("libubsan.so"),
                "FML_ORACLE_LIB": lib})
    probe = ("import oracle; l = oracle.lib(); import sys; "
             "sys.exit(0 if l._name.endswith('liboracle_fml_san.so') else 3)")
    r = subprocess.run([sys.executable, "-c", probe], cwd=ROOT, env=env, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                        os.path.join(ROOT, "tests", "test_oracle.py")],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-3000:])
    assert "passed" in r.stdout


@pytest.mark.skipif(shutil.which("g++") is None or _runtime("libasan.so") is None,
                    reason="needs g++ with ASan and UBSan")
def test_exp_neg_header_under_sanitizers(tmp_path):
    exe = str(tmp_path / "exp_neg_check_san")
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-ffp-contract=off"] + SAN +
                   ["-I", os.path.join(ROOT, "fml_amd", "csrc"),
                    os.path.join(ROOT, "tests", "exp_neg_check.cpp"), "-o", exe], check=True)
    r = subprocess.run([exe, "100000"], capture_output=True, text=True, env=_env(), timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    assert len(r.stdout.splitlines()) == 5
