"""The C ABI: libfmlhip.so loads without a GPU and exports exactly what
include/fmlh.h declares; the Python binding covers every symbol; argument
errors are reported before any device work (no GPU needed)."""
import os
import re
import subprocess

import pytest

import fml_amd._native as nat

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fmlh.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(fmlh_[a-z0-9_]+)\s*\(", text))


def exported_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", nat.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line and "fmlh_" in line}


def test_library_exports_header():
    decl = declared_symbols()
    assert len(decl) >= 20
    exp = exported_symbols()
    missing = decl - exp
    assert not missing, f"declared but not exported: {sorted(missing)}"


def test_binding_covers_header():
    nat.load()
    decl = declared_symbols()
    bound = set(nat.SIGNATURES) | {"fmlh_timing_enable", "fmlh_timing_collect"}
    assert decl <= bound, sorted(decl - bound)


def test_version_and_errors_without_gpu():
    lib = nat.load()
    assert lib.fmlh_abi_version() == 7
    # argument validation happens before any device call
    assert lib.fmlh_potrf(None, 1, -1, None, None, 0, None) == -3
    assert lib.fmlh_gaussian_kernel(None, 0, 5, None, 0, 5, 3, 1.0, None, 5, None) == -2
    assert lib.fmlh_coulomb_sorted(None, None, None, 1, 0, None, None, None, None, None) == -5
    assert lib.fmlh_arad_kernels(9, None, None, None, 1, None, None, None, 1, 5, None, 1, 0.2, 5.0,
                                 1.0, 0.5, 0, None, None, 0, None, None) == -1
    # an out-of-range tiling is rejected before any device work
    bad = nat.AradTiling(16, 4, 16, 0)
    import ctypes
    sig = (ctypes.c_double * 1)(1.0)
    assert lib.fmlh_arad_kernels(0, None, None, None, 1, None, None, None, 1, 5, sig, 1, 0.2, 5.0,
                                 1.0, 0.5, 0, None, None, 0, ctypes.byref(bad), None) == -21
    assert lib.fmlh_arad_tiling_set(0, 23, 0, 3, ctypes.byref(bad)) == -5
    assert lib.fmlh_arad_tiling_set(0, 23, 0, 3, ctypes.byref(nat.AradTiling(512, 5, 8, 2))) == -5
    good = nat.AradTiling(512, 5, 8, 1)
    assert lib.fmlh_arad_tiling_set(0, 23, 1, 3, ctypes.byref(good)) == 0
    got = nat.AradTiling()
    assert lib.fmlh_arad_tiling_get(0, 23, 1, 3, ctypes.byref(got)) == 0
    assert got.as_tuple() == (512, 5, 8, 1)
    assert lib.fmlh_arad_tiling_get(0, 23, 1, 2, ctypes.byref(got)) == 0
    assert got.as_tuple() == (256, 0, 16, 0)  # untuned key: the built-in default
    assert lib.fmlh_graph_launch(None, None) == -1
    assert lib.fmlh_graph_destroy(None) == 0
    describe = lib.fmlh_arad_describe
    assert describe(None, None, None, -1, 23, 5.0, None, 119, None, None, None, None, None) == -4
    assert describe(None, None, None, 1, 0, 5.0, None, 119, None, None, None, None, None) == -5
    assert describe(None, None, None, 1, 65, 5.0, None, 119, None, None, None, None, None) == -5
    assert describe(None, None, None, 1, 23, 0.0, None, 119, None, None, None, None, None) == -6
    assert describe(None, None, None, 1, 23, 5.0, None, 0, None, None, None, None, None) == -8
    assert describe(None, None, None, 1, 23, 5.0, None, 119, None, None, None, None, None) == -1
    assert describe(None, None, None, 0, 23, 5.0, None, 119, None, None, None, None, None) == 0
    # empty problems are no-ops
    assert lib.fmlh_gaussian_kernel(None, 3, 0, None, 3, 0, 3, 1.0, None, 1, None) == 0
    assert lib.fmlh_arad_ws_bytes(10, 10, 23, 0) > 0


def test_no_cpu_fallback(monkeypatch):
    """Without the library the product raises instead of computing on the CPU."""
    monkeypatch.setattr(nat, "_lib", None)
    with pytest.raises(nat.NativeError):
        nat.load("/nonexistent/libfmlhip.so")


def test_predict_workspace_covers_empty_training_set():
    """fmlh_krr_predict takes a kernel block of at least one column, also when
    the training set is empty (its predictions are then zeros): the size
    query must include it."""
    lib = nat.load()
    for nmol in (1, 7, 300):
        assert lib.fmlh_krr_predict_ws_bytes(nmol, 0, 23) == lib.fmlh_krr_predict_ws_bytes(nmol, 1, 23)
        assert lib.fmlh_krr_predict_ws_bytes(nmol, 0, 23) > 0
