"""Molecule container (fml/fml.py:44-132) with GPU descriptor generation."""
from __future__ import annotations

import numpy as np

from .arad import ARAD
from .aras import ARAS
from .data import NUCLEAR_CHARGE
from .representations import (fgenerate_atomic_coulomb_matrix, fgenerate_coulomb_matrix,
                              fgenerate_local_coulomb_matrix, fgenerate_unsorted_coulomb_matrix)


class Molecule:

    def __init__(self):
        self.natoms = -1
        self.energy = float("nan")
        self.molid = -1
        self.name = None
        self.atomtypes = []
        self.nuclear_charges = []
        self.coordinates = []
        self.active_atoms = []
        self.unit_cell = None
        self.properties = []
        self.properties2 = []

    def generate_coulomb_matrix(self, size=23):
        """Sorted Coulomb matrix, fml.py:64-66 (computed on the GPU)."""
        self.coulomb_matrix = fgenerate_coulomb_matrix(self.nuclear_charges, self.coordinates,
                                                       self.natoms, size)

    def generate_unsorted_coulomb_matrix(self, size=23):
        """Unsorted Coulomb matrix, fml.py:68-70 (computed on the GPU)."""
        self.unsorted_coulomb_matrix = fgenerate_unsorted_coulomb_matrix(
            self.nuclear_charges, self.coordinates, self.natoms, size)

    def generate_local_coulomb_matrix(self, calc="all", size=23):
        """Local Coulomb matrices, one row per atom, fml.py:72-74 (GPU)."""
        self.local_coulomb_matrix = fgenerate_local_coulomb_matrix(
            self.nuclear_charges, self.coordinates, self.natoms, size)

    def generate_atomic_coulomb_matrix(self, calc="all", size=23):
        """Atomic Coulomb matrices, one row per atom, fml.py:76-78 (GPU)."""
        self.atomic_coulomb_matrix = fgenerate_atomic_coulomb_matrix(
            self.nuclear_charges, self.coordinates, self.natoms, size)

    def generate_arad_descriptor(self, size=23):
        """ARAD descriptor, fml.py:80-86."""
        arad_object = ARAD(maxMolSize=size, maxAts=size)
        self.arad_descriptor = arad_object.describe(np.array(self.coordinates),
                                                    np.array(self.nuclear_charges))
        assert (self.arad_descriptor).shape[0] == size, "ERROR: Check ARAD descriptor size!"
        assert (self.arad_descriptor).shape[2] == size, "ERROR: Check ARAD descriptor size!"

    def generate_aras_descriptor(self, size=23):
        """ARAS descriptor, fml.py:100-106."""
        aras_object = ARAS(maxMolSize=size, maxAts=size)
        self.aras_descriptor = aras_object.describe(np.array(self.coordinates),
                                                    np.array(self.nuclear_charges))
        assert (self.aras_descriptor).shape[0] == size, "ERROR: Check ARAS descriptor size!"
        assert (self.aras_descriptor).shape[2] == size, "ERROR: Check ARAS descriptor size!"

    def read_xyz(self, filename):
        """XYZ reader, fml.py:109-132: atom count from line 1, atom lines from
        line 3 until a line with fewer than four tokens."""
        with open(filename, "r") as f:
            lines = f.readlines()
        self.natoms = int(lines[0])
        for line in lines[2:]:
            tokens = line.split()
            if len(tokens) < 4:
                break
            self.atomtypes.append(tokens[0])
            self.nuclear_charges.append(NUCLEAR_CHARGE[tokens[0]])
            self.coordinates.append(np.array([float(tokens[1]), float(tokens[2]), float(tokens[3])]))
        self.coordinates = np.array(self.coordinates)
