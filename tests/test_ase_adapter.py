"""The ASE adapter's host logic (no GPU): Voigt order and the sign of a left-handed cell, the
input checks made before the library is touched, and the result cache of the calling convention
(recomputed when positions, numbers or cell change, not when a second property is asked for)."""

import numpy as np
import pytest

import d3_ase
from atoms_like import make_atoms


def test_stress_voigt_order_and_handedness():
    s = np.array([[1.0, 6.0, 5.0], [6.0, 2.0, 4.0], [5.0, 4.0, 3.0]])
    right = np.diag([2.0, 3.0, 4.0])
    np.testing.assert_array_equal(d3_ase.stress_voigt(s, right), [1, 2, 3, 4, 5, 6])
    # the library divides by the signed volume: a left-handed cell gets the sign back
    np.testing.assert_array_equal(d3_ase.stress_voigt(-s, right[[1, 0, 2]]), [1, 2, 3, 4, 5, 6])
    np.testing.assert_array_equal(d3_ase.voigt_to_3x3([1, 2, 3, 4, 5, 6]), s)


@pytest.mark.parametrize("case", ["open", "flat", "empty"])
def test_inputs_refused_before_the_library(case):
    calc = d3_ase.D3ASECalculator()
    cell = np.diag([10.0, 10.0, 10.0])
    if case == "open":
        a = make_atoms([1, 1], [[0, 0, 0], [0.7, 0, 0]], cell, pbc=False)
    elif case == "flat":
        a = make_atoms([1, 1], [[0, 0, 0], [0.7, 0, 0]], np.diag([10.0, 10.0, 0.0]))
    else:
        a = make_atoms([], np.zeros((0, 3)), cell)
    with pytest.raises(ValueError):
        calc.get_potential_energy(a)
    assert calc._handle is None


@pytest.mark.skipif(d3_ase.HAVE_ASE, reason="checks the stand-in base used when ASE is absent")
def test_result_cache_of_the_calling_convention():
    class Counting(d3_ase.D3ASECalculator):
        def calculate(self, atoms=None, properties=None, system_changes=d3_ase.all_changes):
            d3_ase._Base.calculate(self, atoms, properties, system_changes)
            self.evaluations += 1
            n = len(self.atoms.get_atomic_numbers())
            self.results = {"energy": 1.0, "free_energy": 1.0, "forces": np.zeros((n, 3)),
                            "stress": np.zeros(6)}

    calc = Counting()
    a = make_atoms([8, 1, 1], [[0, 0, 0], [0.96, 0, 0], [-0.24, 0.93, 0]], np.diag([12.0, 12.0, 12.0]))
    calc.get_potential_energy(a)
    calc.get_forces(a)
    calc.get_stress(a)
    assert calc.evaluations == 1
    p = a.get_positions()
    p[0, 0] += 0.01
    a.set_positions(p)
    calc.get_forces(a)
    assert calc.evaluations == 2
    assert calc.get_stress(a, voigt=False).shape == (3, 3) and calc.evaluations == 2
