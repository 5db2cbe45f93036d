"""Persistence of ARAD tile-shape tuning results (no reference counterpart).

``AradKernelCall.autotune`` (fml_amd.batch) times candidate tilings on the
device and stores the winner in libfmlhip's process-wide table, keyed by
(mode, amax, precision, nsig); calls without an explicit tiling use it.
This module keeps a record of those results so they can be written to a
JSON file and loaded into the table of a later process:

    tuning.save("arad_tiling.json")      # after autotuning
    tuning.load("arad_tiling.json")      # in a new process
"""
from __future__ import annotations

import ctypes
import json

from . import _native

#: (mode, amax, precision, nsig) -> (pairs_per_tile, waves_per_simd, groups_per_cu, pair_order)
RESULTS: dict = {}


def record(mode: int, amax: int, precision: int, nsig: int, tiling) -> None:
    RESULTS[(int(mode), int(amax), int(precision), int(nsig))] = tuple(tiling.as_tuple())


def current(mode: int, amax: int, precision: int, nsig: int):
    """The library's tiling for this key (tuned or default)."""
    t = _native.AradTiling()
    _native.call("fmlh_arad_tiling_get", int(mode), int(amax), int(precision), int(nsig), ctypes.byref(t))
    return t


def save(path: str) -> None:
    rows = [{"mode": k[0], "amax": k[1], "precision": k[2], "nsig": k[3],
             "pairs_per_tile": v[0], "waves_per_simd": v[1], "groups_per_cu": v[2],
             "pair_order": v[3]}
            for k, v in sorted(RESULTS.items())]
    with open(path, "w") as f:
        json.dump(rows, f, indent=1)


def load(path: str) -> int:
    """Load saved results into the library's table; returns the count."""
    with open(path) as f:
        rows = json.load(f)
    for r in rows:
        t = _native.AradTiling(int(r["pairs_per_tile"]), int(r["waves_per_simd"]), int(r["groups_per_cu"]),
                               int(r.get("pair_order", 0)))
        _native.call("fmlh_arad_tiling_set", int(r["mode"]), int(r["amax"]), int(r["precision"]),
                     int(r["nsig"]), ctypes.byref(t))
        record(r["mode"], r["amax"], r["precision"], r["nsig"], t)
    return len(rows)
