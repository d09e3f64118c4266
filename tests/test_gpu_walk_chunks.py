"""Pair walks at chunk boundaries: a simple cubic lattice whose bins hold exactly 64 atoms (one
full chunk of candidates) or exactly 125 (a full chunk and a partial one), with the lattice
points moved within their bins, through the CN, two-body and chain-rule walks (the FP32
kernels load each chunk one ahead, enumerating chunks across a group's bins by a lane scan; the
FP64 two-body kernel walks bin by bin), against the oracle: FP64 at 1e-12 (energy, CN) and 1e-9
eV/A (forces), FP32 at its 2e-6 / 2e-5 bar."""

import numpy as np
import pytest

import d3_lib
import d3_oracle

pytestmark = pytest.mark.gpu

BOHR_A = 0.52917726
S_A = 2.5  # lattice spacing, Angstrom
N = 20     # lattice points per side


def lattice():
    # points at the middle of their lattice cells, moved by at most 0.4 A (the bins' walls are
    # 1.25 A away), so every bin keeps its exact count
    g = np.stack(np.meshgrid(*[np.arange(N)] * 3, indexing="ij"), -1).reshape(-1, 3)
    rng = np.random.Generator(np.random.PCG64(17))
    pos = (g + 0.5) * S_A + np.clip(rng.normal(0.0, 0.1, (len(g), 3)), -0.4, 0.4)
    return pos, np.full(len(pos), 14, dtype=np.uint16), np.eye(3) * (N * S_A)


@pytest.mark.parametrize("per_side", [4, 5])
@pytest.mark.parametrize("precision", [d3_lib.D3X_FP64, d3_lib.D3X_FP32])
def test_walk_at_chunk_boundaries(table, per_side, precision):
    pos, Z, cell = lattice()
    kw = dict(cutoff=45.0, cn_cutoff=30.0, atm_cutoff=0.0, damping=1, functional=0)
    print(f"case walk chunks {per_side ** 3} atoms per bin precision={precision}", flush=True)
    h = d3_lib.Handle(np.unique(Z), len(Z), precision=precision, **kw)
    # bins of per_side lattice spacings (slightly under, so that the cell holds N / per_side bins)
    h.set_tuning(bin_width=per_side * S_A / BOHR_A * (1 - 1e-4))
    h.set_atoms(pos, Z)
    h.set_cell(cell)
    e, f, st = h.compute()
    cn, _, _ = h.atom_data()
    assert abs(h.tuning()[0] - per_side * S_A / BOHR_A * (1 - 1e-4)) < 1e-9
    h.close()
    o = d3_oracle.compute(pos, Z, cell, table, **kw)
    if precision == d3_lib.D3X_FP64:
        assert abs(e - o["energy"]) <= 1e-12 * abs(o["energy"])
        np.testing.assert_allclose(cn, o["cn"], rtol=1e-12, atol=0)
        assert np.max(np.abs(f - o["forces"])) <= 1e-9
        assert np.max(np.abs(st - o["stress"])) <= 1e-12 * np.max(np.abs(o["stress"])) + 1e-12
    else:
        assert abs(e - o["energy"]) <= 2e-6 * abs(o["energy"])
        np.testing.assert_allclose(cn, o["cn"], rtol=2e-6, atol=0)
        assert np.max(np.abs(f - o["forces"])) <= 2e-5 * max(1e-3, np.max(np.abs(o["forces"])))
