"""Sorted Coulomb matrices on the GPU.

Drop-in for fml.representations.fgenerate_coulomb_matrix
(fml/representations/frepresentations.f90:23-110, exported at
fml/representations/__init__.py:31) plus a batched form.
"""
from __future__ import annotations

import numpy as np

from .. import _native
from ..device import check_offsets, empty, is_device_tensor, to_device, to_host


def generate_coulomb_matrices(charges, coords, offsets, nmax: int = 23, float32: bool = False,
                              return_perm: bool = False, on_device: bool = False):
    """Sorted Coulomb matrices of many molecules in one launch.

    charges (total,), coords (total, 3), offsets (nmol+1,) int: molecule m
    owns atoms offsets[m]:offsets[m+1].  Returns (nmol, nmax(nmax+1)/2)
    float64 (and the float32 copy when float32=True), optionally the
    (nmol, nmax) int32 sorted-position -> atom permutation.  Raises
    ValueError when a molecule has more than nmax atoms (the reference
    writes out of bounds there, frepresentations.f90:96-103).
    """
    z = to_device(charges, np.float64)
    r = to_device(np.asarray(coords).reshape(-1, 3) if not is_device_tensor(coords) else coords,
                  np.float64)
    if int(r.shape[0]) != int(z.shape[0]):
        raise ValueError(f"{int(r.shape[0])} coordinates for {int(z.shape[0])} charges")
    check_offsets(offsets, int(z.shape[0]))
    off = to_device(offsets, np.int32)
    nmol = int(off.shape[0]) - 1
    packed = nmax * (nmax + 1) // 2
    cm = empty((nmol, packed), np.float64)
    cm32 = empty((nmol, packed), np.float32) if float32 else None
    perm = empty((nmol, nmax), np.int32) if return_perm else None
    status = empty((max(nmol, 1),), np.int32)
    _native.call("fmlh_coulomb_sorted", _native.ptr(z), _native.ptr(r), _native.ptr(off), nmol, nmax,
                 _native.ptr(cm), _native.ptr(cm32), _native.ptr(perm), _native.ptr(status),
                 _native.stream_ptr())
    if nmol and int(status[:nmol].max().item()) != 0:
        bad = int(np.nonzero(to_host(status[:nmol]))[0][0])
        raise ValueError(f"molecule {bad} has more than nmax={nmax} atoms")
    outs = [cm]
    if float32:
        outs.append(cm32)
    if return_perm:
        outs.append(perm)
    if not on_device:
        outs = [to_host(o) for o in outs]
    return outs[0] if len(outs) == 1 else tuple(outs)


def fgenerate_coulomb_matrix(nuclear_charges, coordinates, natoms, nmax):
    """frepresentations.f90:23 -- one molecule -> (nmax(nmax+1)/2,) float64.

    The reference stops the interpreter when len(coordinates) differs from
    len(nuclear_charges) (frepresentations.f90:45-49); here that is a
    ValueError.  natoms is ignored as in the reference (it is recomputed from
    the charges, line 51)."""
    z = np.asarray(nuclear_charges, dtype=np.float64).reshape(-1)
    r = np.asarray(coordinates, dtype=np.float64).reshape(-1, 3)
    if r.shape[0] != z.shape[0]:
        raise ValueError(f"ERROR: Coulomb matrix generation: {r.shape[0]} coordinates, "
                         f"but {z.shape[0]} atom_types!")
    off = np.array([0, z.shape[0]], dtype=np.int32)
    return generate_coulomb_matrices(z, r, off, int(nmax))[0]


UNSORTED, LOCAL, ATOMIC = 0, 1, 2


def _coulomb_form(mode, charges, coords, offsets, nmax, on_device):
    z = to_device(charges, np.float64)
    r = to_device(np.asarray(coords).reshape(-1, 3) if not is_device_tensor(coords) else coords,
                  np.float64)
    if int(r.shape[0]) != int(z.shape[0]):
        raise ValueError(f"{int(r.shape[0])} coordinates for {int(z.shape[0])} charges")
    check_offsets(offsets, int(z.shape[0]))
    off = to_device(offsets, np.int32)
    nmol = int(off.shape[0]) - 1
    packed = nmax * (nmax + 1) // 2
    rows = nmol if mode == UNSORTED else int(z.shape[0])
    cm = empty((rows, packed), np.float64)
    status = empty((max(nmol, 1),), np.int32)
    _native.call("fmlh_coulomb_form", mode, _native.ptr(z), _native.ptr(r), _native.ptr(off), nmol,
                 int(nmax), _native.ptr(cm), _native.ptr(status), _native.stream_ptr())
    if nmol and int(status[:nmol].max().item()) != 0:
        bad = int(np.nonzero(to_host(status[:nmol]))[0][0])
        raise ValueError(f"molecule {bad} has more than nmax={nmax} atoms")
    return cm if on_device else to_host(cm)


def generate_unsorted_coulomb_matrices(charges, coords, offsets, nmax: int = 23,
                                       on_device: bool = False):
    """Unsorted Coulomb matrices of many molecules in one launch
    (fgenerate_unsorted_coulomb_matrix, frepresentations.f90:112-201):
    (nmol, nmax(nmax+1)/2), the packed lower triangle in input atom order."""
    return _coulomb_form(UNSORTED, charges, coords, offsets, nmax, on_device)


def generate_local_coulomb_matrices(charges, coords, offsets, nmax: int = 23,
                                    on_device: bool = False):
    """Local Coulomb matrices of many molecules in one launch
    (fgenerate_local_coulomb_matrix, frepresentations.f90:204-294):
    (total atoms, nmax(nmax+1)/2), row offsets[m] + k for atom k of molecule
    m: atom k first, then the other atoms by descending row norm."""
    return _coulomb_form(LOCAL, charges, coords, offsets, nmax, on_device)


def generate_atomic_coulomb_matrices(charges, coords, offsets, nmax: int = 23,
                                     on_device: bool = False):
    """Atomic Coulomb matrices of many molecules in one launch
    (fgenerate_atomic_coulomb_matrix, frepresentations.f90:297-390):
    (total atoms, nmax(nmax+1)/2), row offsets[m] + k ordering the atoms by
    ascending distance from atom k (ties lowest index first)."""
    return _coulomb_form(ATOMIC, charges, coords, offsets, nmax, on_device)


def _one(mode, nuclear_charges, coordinates, nmax):
    z = np.asarray(nuclear_charges, dtype=np.float64).reshape(-1)
    r = np.asarray(coordinates, dtype=np.float64).reshape(-1, 3)
    if r.shape[0] != z.shape[0]:
        raise ValueError(f"ERROR: Coulomb matrix generation: {r.shape[0]} coordinates, "
                         f"but {z.shape[0]} atom_types!")
    off = np.array([0, z.shape[0]], dtype=np.int32)
    return _coulomb_form(mode, z, r, off, int(nmax), False)


def fgenerate_unsorted_coulomb_matrix(nuclear_charges, coordinates, natoms, nmax):
    """frepresentations.f90:112 -- one molecule -> (nmax(nmax+1)/2,).  natoms
    is recomputed from the charges, as in the reference (:134-141)."""
    return _one(UNSORTED, nuclear_charges, coordinates, nmax)[0]


def fgenerate_local_coulomb_matrix(nuclear_charges, coordinates, natoms, nmax):
    """frepresentations.f90:204 -- one molecule -> (natoms, nmax(nmax+1)/2).
    A length mismatch raises ValueError (the reference stops, :226-232); more
    than nmax atoms raises (the reference writes out of bounds)."""
    return _one(LOCAL, nuclear_charges, coordinates, nmax)


def fgenerate_atomic_coulomb_matrix(nuclear_charges, coordinates, natoms, nmax):
    """frepresentations.f90:297 -- one molecule -> (natoms, nmax(nmax+1)/2)."""
    return _one(ATOMIC, nuclear_charges, coordinates, nmax)


__all__ = ["fgenerate_coulomb_matrix", "generate_coulomb_matrices",
           "fgenerate_unsorted_coulomb_matrix", "fgenerate_local_coulomb_matrix",
           "fgenerate_atomic_coulomb_matrix", "generate_unsorted_coulomb_matrices",
           "generate_local_coulomb_matrices", "generate_atomic_coulomb_matrices"]
