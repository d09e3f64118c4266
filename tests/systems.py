"""Test systems: the reference's six gtest systems (test/d3_test.cpp:343-431) and small periodic
cells that exercise image handling, triclinic cells, mixed elements and the three-body term."""

from __future__ import annotations

import numpy as np

import d3_structures as S

SQRT9000 = 94.86832980505137

# test/d3_test.cpp:398-417: C3OH6 molecule in a 20 A cube, cutoffs 40/40 bohr, ZERO, PBE0
SMALL = dict(
    name="SmallSystem",
    Z=np.array([6, 6, 6, 8, 1, 1, 1, 1, 1, 1], np.uint16),
    pos=np.array([[5.1372, 5.5512, 10.1047], [4.5169, 6.1365, 11.3604], [6.1937, 4.4752, 10.2703],
                  [4.7872, 5.9358, 8.9937], [6.7474, 4.3475, 9.3339], [5.6975, 3.5214, 10.5181],
                  [6.8870, 4.7006, 11.0939], [4.8579, 5.6442, 12.2774], [3.4204, 6.0677, 11.2935],
                  [4.7678, 7.2075, 11.4098]], np.float32).astype(np.float64),
    cell=np.eye(3) * 20.0, cutoff=40.0, cn_cutoff=40.0, damping=0, functional=1)


def bcc_fe(surface=False):
    """test/d3_test.cpp:343-396: BCC Fe 2x2x2, a = 2.87 A (surface: c x 1.2)."""
    a = np.float32(2.87)
    pos = []
    for i in range(2):
        for j in range(2):
            for k in range(2):
                pos.append([i * a, j * a, k * a])
                pos.append([(i + np.float32(0.5)) * a, (j + np.float32(0.5)) * a, (k + np.float32(0.5)) * a])
    cell = np.diag(np.array([2 * a, 2 * a, 2 * a], np.float32))
    if surface:
        cell[2, 2] = cell[2, 2] * np.float32(1.2)
    return dict(name="BCC_Fe_Surface" if surface else "BCC_Fe_Crystal", Z=np.full(16, 26, np.uint16),
                pos=np.array(pos, np.float32).astype(np.float64), cell=cell.astype(np.float64), cutoff=40.0,
                cn_cutoff=40.0, damping=0, functional=1)


def reference_systems():
    out = []
    for s in (SMALL, bcc_fe(False), bcc_fe(True)):
        out.append(dict(s))
        bj = dict(s)
        bj["name"] = s["name"] + "_BJ"
        bj["damping"] = 1
        out.append(bj)
    return out


def si64(seed=1):
    pos, Z, cell = S.si_diamond(nrep=2, jitter=0.02, seed=seed)
    return dict(name="Si64", Z=Z, pos=pos, cell=cell, cutoff=SQRT9000, cn_cutoff=40.0, atm=40.0, damping=1,
                functional=0)


def triclinic_mixed(n=24, seed=3):
    rng = np.random.Generator(np.random.PCG64(seed))
    cell = S.triclinic_cell(n / 0.08, rng.uniform(70, 110, 3))
    pos, Z = S.random_packing(n, cell, rng)
    return dict(name="Triclinic6", Z=Z, pos=pos, cell=cell, cutoff=50.0, cn_cutoff=30.0, atm=30.0, damping=1,
                functional=2)


def water_small(n_mol=16, seed=5):
    pos, Z, cell = S.water_box(n_mol=n_mol, seed=seed)
    return dict(name=f"Water{n_mol}", Z=Z, pos=pos, cell=cell, cutoff=SQRT9000, cn_cutoff=40.0, atm=30.0,
                damping=1, functional=0)


def skewed_zero(seed=7):
    """Strongly skewed cell, zero damping, two elements, ATM with the image range set by rc3."""
    rng = np.random.Generator(np.random.PCG64(seed))
    cell = np.array([[9.0, 0.0, 0.0], [6.5, 8.0, 0.0], [-3.0, 2.5, 7.5]])
    pos, _ = S.random_packing(12, cell, rng)
    Z = np.array([8, 14] * 6, np.uint16)
    return dict(name="SkewedZero", Z=Z, pos=pos, cell=cell, cutoff=35.0, cn_cutoff=25.0, atm=45.0, damping=0,
                functional=5)


def parity_systems():
    return reference_systems() + [si64(), triclinic_mixed(), water_small(), skewed_zero()]
