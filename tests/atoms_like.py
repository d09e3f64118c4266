"""A stand-in for ase.Atoms (ASE is not installed in this image): the methods and attribute the
ASE adapter reads.  When ASE is importable the tests use ase.Atoms itself (make_atoms)."""

import numpy as np


class AtomsLike:
    def __init__(self, numbers, positions, cell, pbc=True):
        self.numbers = np.array(numbers, dtype=int)
        self.positions = np.array(positions, dtype=float).reshape(-1, 3)
        self.cell = np.array(cell, dtype=float).reshape(3, 3)
        self.pbc = np.array([pbc] * 3 if np.isscalar(pbc) else pbc, dtype=bool)

    def get_atomic_numbers(self):
        return self.numbers.copy()

    def get_positions(self):
        return self.positions.copy()

    def set_positions(self, p):
        self.positions = np.array(p, dtype=float).reshape(-1, 3)

    def get_cell(self):
        return self.cell.copy()

    def get_volume(self):
        return abs(np.linalg.det(self.cell))

    def copy(self):
        return AtomsLike(self.numbers, self.positions, self.cell, self.pbc)

    def __len__(self):
        return len(self.numbers)


def make_atoms(numbers, positions, cell, pbc=True):
    try:
        import ase
        return ase.Atoms(numbers=numbers, positions=positions, cell=cell, pbc=pbc)
    except ImportError:
        return AtomsLike(numbers, positions, cell, pbc)
