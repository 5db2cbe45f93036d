"""Seeded synthetic molecule sets for the benchmark and parity workloads.

Atom counts are uniform on [nmin, nmax] (or truncated geometric), elements
H/C/N/O/S with probabilities .5/.3/.1/.08/.02, coordinates rejection-sampled
atom by atom in a cube of edge 1.2*n^(1/3) Angstrom with every pair at least
0.9 Angstrom apart (a molecule restarts after `max_tries` failed draws for
one atom), targets = sum of per-element constants + N(0, 0.01^2).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

ELEMENTS = np.array([1, 6, 7, 8, 16], dtype=np.int64)
ELEMENT_P = np.array([0.5, 0.3, 0.1, 0.08, 0.02])
ELEMENT_ENERGY = {1: -0.5, 6: -38.0, 7: -54.6, 8: -75.0, 16: -398.0}


@dataclass
class MoleculeSet:
    """Concatenated molecules: charges[offsets[m]:offsets[m+1]] belong to
    molecule m, coords likewise (rows of xyz)."""
    charges: np.ndarray   # (total,) int64
    coords: np.ndarray    # (total, 3) float64
    offsets: np.ndarray   # (nmol+1,) int32
    energies: np.ndarray  # (nmol,) float64

    @property
    def nmol(self) -> int:
        return self.offsets.shape[0] - 1

    @property
    def natoms(self) -> np.ndarray:
        return np.diff(self.offsets)

    def molecule(self, m: int):
        a, b = self.offsets[m], self.offsets[m + 1]
        return self.charges[a:b], self.coords[a:b]

    def subset(self, idx) -> "MoleculeSet":
        idx = np.asarray(idx)
        zs, rs, ns = [], [], []
        for m in idx:
            z, r = self.molecule(int(m))
            zs.append(z)
            rs.append(r)
            ns.append(len(z))
        off = np.zeros(len(idx) + 1, dtype=np.int32)
        off[1:] = np.cumsum(ns)
        return MoleculeSet(np.concatenate(zs) if zs else np.zeros(0, np.int64),
                           np.concatenate(rs) if rs else np.zeros((0, 3)),
                           off, self.energies[idx].copy())


def _place(n: int, rng: np.random.Generator, min_dist: float, max_tries: int) -> np.ndarray:
    # the minimum-distance test in plain floats: (dx*dx + dy*dy) + dz*dz is
    # exactly numpy's sum((a - b)**2) over three values
    edge = 1.2 * n ** (1.0 / 3.0)
    md2 = min_dist * min_dist
    while True:
        pts = []
        for k in range(n):
            for _ in range(max_tries):
                x, y, z = rng.uniform(0.0, edge, 3).tolist()
                if all((x - a) * (x - a) + (y - b) * (y - b) + (z - c) * (z - c) >= md2
                       for a, b, c in pts):
                    pts.append((x, y, z))
                    break
            else:
                break
        if len(pts) == n:
            return np.array(pts, dtype=np.float64).reshape(n, 3)


def atom_counts(nmol: int, rng: np.random.Generator, nmin: int, nmax: int,
                dist: str = "uniform", p_geom: float = 0.0329) -> np.ndarray:
    if dist == "uniform":
        return rng.integers(nmin, nmax + 1, size=nmol)
    if dist == "geometric":
        # truncated geometric on [nmin, nmax]: P(n) ~ (1-p)^(n-nmin)
        ks = np.arange(nmin, nmax + 1)
        w = (1.0 - p_geom) ** (ks - nmin)
        return rng.choice(ks, size=nmol, p=w / w.sum())
    raise ValueError(dist)


def generate(nmol: int, seed: int = 0, nmin: int = 5, nmax: int = 23, dist: str = "uniform",
             min_dist: float = 0.9, max_tries: int = 2000) -> MoleculeSet:
    rng = np.random.default_rng(seed)
    counts = atom_counts(nmol, rng, nmin, nmax, dist)
    zs, rs = [], []
    energies = np.empty(nmol)
    for m, n in enumerate(counts):
        z = rng.choice(ELEMENTS, size=int(n), p=ELEMENT_P)
        r = _place(int(n), rng, min_dist, max_tries)
        zs.append(z)
        rs.append(r)
        energies[m] = sum(ELEMENT_ENERGY[int(e)] for e in z)
    energies += rng.normal(0.0, 0.01, size=nmol)
    off = np.zeros(nmol + 1, dtype=np.int32)
    off[1:] = np.cumsum(counts)
    return MoleculeSet(np.concatenate(zs), np.concatenate(rs), off, energies)


def save(ms: MoleculeSet, path: str):
    np.savez_compressed(path, charges=ms.charges, coords=ms.coords, offsets=ms.offsets,
                        energies=ms.energies)


def load(path: str) -> MoleculeSet:
    d = np.load(path)
    return MoleculeSet(d["charges"], d["coords"], d["offsets"], d["energies"])
