"""Build libfmlhip.so (gfx950) in-tree with hipcc.

    python -m fml_amd.build          # or fml_amd.build.build()

Every HIP source under fml_amd/csrc is compiled for gfx950 only (one hipcc
process per source, in parallel) and linked into fml_amd/libfmlhip.so next
to this file.  Packed FP32 VALU instructions are disabled for device code
(-target-feature -packed-fp32-ops).
"""
from __future__ import annotations

import glob
import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libfmlhip.so")
ARCH = "gfx950"


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def sources():
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))


def _jobs() -> int:
    try:
        return max(1, min(len(sources()), int(os.environ.get("MAX_JOBS", "0")) or os.cpu_count() or 1, 16))
    except ValueError:
        return 4


def build(verbose: bool = False, extra_flags=None) -> str:
    """Compile every source into LIB (always recompiles).  Returns LIB."""
    flags = [f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
             "-I", os.path.join(ROOT, "include"), "-I", CSRC,
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument"] + (extra_flags or [])
    with tempfile.TemporaryDirectory(prefix="fmlh_build_") as tmpd:
        def compile_one(src):
            obj = os.path.join(tmpd, os.path.basename(src) + ".o")
            cmd = [hipcc()] + flags + ["-c", src, "-o", obj]
            if verbose:
                print(" ".join(cmd), flush=True)
            r = subprocess.run(cmd, capture_output=True, text=True)
            return src, obj, r
        with ThreadPoolExecutor(max_workers=_jobs()) as ex:
            results = list(ex.map(compile_one, sources()))
        failed = [(s, r) for s, _, r in results if r.returncode != 0]
        if failed:
            for s, r in failed:
                sys.stderr.write(f"---- {s}\n" + r.stdout + r.stderr)
            raise RuntimeError("hipcc failed building libfmlhip.so")
        tmp = LIB + ".tmp"
        cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + [o for _, o, _ in results]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("hipcc failed linking libfmlhip.so")
        os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
