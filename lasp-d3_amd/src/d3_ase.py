"""ASE calculator over the D3 engine: energy, forces and stress of an ``ase.Atoms``.

The reference's ASE example (examples/ase_mace_laspd3.py:17-100) wraps its D3Calculator in an
ASE ``Calculator`` that adds D3 to a MACE model.  ``D3ASECalculator`` is that adapter for the D3
part: it returns the dispersion energy (eV), forces (eV/A) and stress (eV/A^3, ASE's Voigt order
xx, yy, zz, yz, xz, xy), and adds the results of an optional ``base`` calculator -- the MACE model
of the example, or any ASE calculator -- the way the example's ``calculate`` sums the two.

Where it departs from the example, on purpose:

* the stress is not divided by the volume again: the library's stress is already per volume
  (src/d3.cu:210-218 converts hartree/bohr^3 to eV/A^3), so the example's division
  (ase_mace_laspd3.py:110-119) makes it wrong by a factor of V;
* the library's stress is (1/V) dE/d(strain) with V the signed determinant of the cell
  (src/d3_kernel.cu:90-96), so it changes sign for a left-handed cell; ASE's stress is the
  physical one, so the adapter multiplies by the sign of the determinant;
* positions, cell and results travel in double precision (the d3x entry points), and
  ``precision`` chooses the engine's FP32 compensated path (default, as the reference's float
  entry points) or its FP64 path;
* the handle is kept across calls and rebuilt only when the atoms bring an element it was not
  made for or more atoms than it holds; ``skin`` (Angstrom) keeps the neighbour state across MD
  steps while no atom has moved more than half of it.

ASE itself is optional: when it is importable the class derives from
``ase.calculators.calculator.Calculator``; otherwise from a small base with the same calling
convention (``calculate(atoms, properties, system_changes)``, ``results``, ``get_property``,
``get_potential_energy``, ``get_forces``, ``get_stress``) that accepts any object with
``get_atomic_numbers()``, ``get_positions()``, ``get_cell()``, ``pbc`` and ``copy()``.
"""

from __future__ import annotations

import numpy as np

import d3_lib

BOHR_A = 0.52917726  # Angstrom per bohr, the value the library converts with (src/d3.cu:50)

try:
    from ase.calculators.calculator import Calculator as _Base
    from ase.calculators.calculator import all_changes
    HAVE_ASE = True
except ImportError:  # ASE absent: the same calling convention on a minimal base
    HAVE_ASE = False
    all_changes = ["positions", "numbers", "cell", "pbc"]

    class _Base:
        implemented_properties: list = []

        def __init__(self, **kwargs):
            self.atoms = None
            self.results = {}
            self.parameters = dict(kwargs)

        @staticmethod
        def _state(atoms):
            return (np.array(atoms.get_atomic_numbers()), np.array(atoms.get_positions(), dtype=float),
                    np.array(atoms.get_cell(), dtype=float).reshape(3, 3), np.array(atoms.pbc, dtype=bool))

        def check_state(self, atoms):
            if self.atoms is None:
                return list(all_changes)
            a, b = self._state(self.atoms), self._state(atoms)
            return [name for name, x, y in zip(("numbers", "positions", "cell", "pbc"), a, b)
                    if x.shape != y.shape or not np.array_equal(x, y)]

        def calculate(self, atoms=None, properties=None, system_changes=all_changes):
            if atoms is not None:
                self.atoms = atoms.copy()

        def get_property(self, name, atoms=None, allow_calculation=True):
            if name not in self.implemented_properties:
                raise NotImplementedError(f"{name} property not implemented")
            if atoms is None:
                atoms = self.atoms
                changes = []
            else:
                changes = self.check_state(atoms)
                if changes:
                    self.results = {}
            if name not in self.results:
                if not allow_calculation:
                    return None
                self.calculate(atoms, [name], changes or list(all_changes))
            r = self.results[name]
            return r.copy() if isinstance(r, np.ndarray) else r

        def get_potential_energy(self, atoms=None, force_consistent=False):
            return self.get_property("free_energy" if force_consistent else "energy", atoms)

        def get_forces(self, atoms=None):
            return self.get_property("forces", atoms)

        def get_stress(self, atoms=None, voigt=True):
            s = self.get_property("stress", atoms)
            return s if voigt else voigt_to_3x3(s)


def stress_voigt(stress_3x3, cell) -> np.ndarray:
    """ASE's Voigt stress (xx, yy, zz, yz, xz, xy) from the library's 3x3 stress: per volume
    already; times the sign of det(cell), since the library divides by the signed volume."""
    s = np.asarray(stress_3x3, dtype=float).reshape(3, 3)
    sign = 1.0 if np.linalg.det(np.asarray(cell, dtype=float).reshape(3, 3)) > 0 else -1.0
    return sign * np.array([s[0, 0], s[1, 1], s[2, 2], s[1, 2], s[0, 2], s[0, 1]])


def voigt_to_3x3(v) -> np.ndarray:
    xx, yy, zz, yz, xz, xy = np.asarray(v, dtype=float)
    return np.array([[xx, xy, xz], [xy, yy, yz], [xz, yz, zz]])


class D3ASECalculator(_Base):
    """DFT-D3 dispersion (two-body, optionally three-body) as an ASE calculator, plus an
    optional ``base`` calculator whose energy, forces and stress are added.

    Cutoffs are in bohr, as in the reference's D3Calculator (src/d3_cffi.py:51-53):
    ``cutoff_radius`` (two-body), ``cn_cutoff_radius`` (coordination number), ``atm_cutoff``
    (three-body; None leaves the term off, as the reference's ``init_d3_handle`` does).
    ``damping_type`` 0 = zero damping, 1 = Becke-Johnson; ``functional_type`` indexes the
    reference's functional table (0 = PBE).  ``precision`` is "fp32" or "fp64"."""

    implemented_properties = ["energy", "free_energy", "forces", "stress"]

    def __init__(self, cutoff_radius: float = 46.4758, cn_cutoff_radius: float = 46.4758,
                 atm_cutoff: float | None = None, damping_type: int = 0, functional_type: int = 0,
                 precision: str = "fp32", max_length: int = 0, skin: float = 0.0, base=None, **kwargs):
        super().__init__(**kwargs)
        if precision not in ("fp32", "fp64"):
            raise ValueError("precision must be 'fp32' or 'fp64'")
        self.cutoff_radius = float(cutoff_radius)
        self.cn_cutoff_radius = float(cn_cutoff_radius)
        self.atm_cutoff = 0.0 if atm_cutoff is None else float(atm_cutoff)
        self.damping_type = int(damping_type)
        self.functional_type = int(functional_type)
        self.precision = d3_lib.D3X_FP32 if precision == "fp32" else d3_lib.D3X_FP64
        self.max_length = int(max_length)
        self.skin = float(skin)
        self.base = base
        self._handle = None
        self._elements = None
        self._capacity = 0
        self.evaluations = 0

    def _get_handle(self, Z: np.ndarray):
        el = np.unique(Z)
        fits = (self._handle is not None and len(Z) <= self._capacity
                and np.isin(el, self._elements).all())
        if not fits:
            if self._handle is not None:
                self._handle.close()
                self._handle = None
            cap = max(self.max_length, len(Z))
            self._handle = d3_lib.Handle(el, cap, cutoff=self.cutoff_radius, cn_cutoff=self.cn_cutoff_radius,
                                         atm_cutoff=self.atm_cutoff, damping=self.damping_type,
                                         functional=self.functional_type, precision=self.precision)
            if self.skin > 0:
                self._handle.set_skin(self.skin / BOHR_A)
            self._elements, self._capacity = el, cap
        return self._handle

    def calculate(self, atoms=None, properties=None, system_changes=all_changes):
        if properties is None:
            properties = self.implemented_properties
        super().calculate(atoms, properties, system_changes)
        a = self.atoms
        Z = np.ascontiguousarray(a.get_atomic_numbers(), dtype=np.uint16)
        pos = np.ascontiguousarray(a.get_positions(), dtype=np.float64)
        cell = np.ascontiguousarray(np.asarray(a.get_cell(), dtype=np.float64).reshape(3, 3))
        # checked before the library is touched
        if not np.all(np.asarray(a.pbc, dtype=bool)):
            raise ValueError("the D3 engine is periodic in all three directions: set pbc=True "
                             "(a molecule goes into a box wider than the cutoffs)")
        if not abs(np.linalg.det(cell)) > 0:
            raise ValueError("the cell has zero volume")
        if len(Z) == 0:
            raise ValueError("no atoms")
        h = self._get_handle(Z)
        h.set_atoms(pos, Z)
        h.set_cell(cell)
        e, f, s = h.compute()
        self.evaluations += 1
        res = {"energy": e, "free_energy": e, "forces": f, "stress": stress_voigt(s, cell)}
        if self.base is not None:
            self.base.calculate(a, properties, system_changes)
            b = self.base.results
            res["energy"] = res["energy"] + b["energy"]
            res["free_energy"] = res["free_energy"] + b.get("free_energy", b["energy"])
            res["forces"] = res["forces"] + np.asarray(b["forces"])
            res["stress"] = res["stress"] + np.asarray(b.get("stress", np.zeros(6)))
        self.results = res

    def close(self):
        if self._handle is not None:
            self._handle.close()
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
