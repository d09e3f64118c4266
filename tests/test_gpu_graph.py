"""hipGraph mode: every replay of the captured evaluation equals the eager evaluation, with eager
work and synchronisations between replays; new coordinates with the same atom count replay the
same graph; a changed cell or atom count re-captures."""

import numpy as np
import pytest

import d3_lib
import d3_structures as S
import systems

pytestmark = pytest.mark.gpu

KW = dict(cutoff=60.0, cn_cutoff=30.0, atm_cutoff=30.0, damping=1, functional=0)


def eager(pos, Z, cell, precision):
    h = d3_lib.Handle(np.unique(Z), len(Z), precision=precision, **KW)
    h.set_atoms(pos, Z)
    h.set_cell(cell)
    out = h.compute()
    h.close()
    return out


@pytest.mark.parametrize("precision", [d3_lib.D3X_FP64, d3_lib.D3X_FP32])
def test_graph_replays_match_eager(precision):
    s = systems.water_small(n_mol=32, seed=3)
    pos, Z, cell = s["pos"], s["Z"], s["cell"]
    rng = np.random.Generator(np.random.PCG64(5))
    hg = d3_lib.Handle(np.unique(Z), len(Z), precision=precision, **KW)
    hg.set_graph(True)
    hg.set_cell(cell)
    for step in range(6):
        p = pos + rng.normal(0.0, 0.05, pos.shape)
        hg.set_atoms(p, Z)
        eg, fg, sg = hg.compute()               # capture on step 0, replay afterwards
        ee, fe, se = eager(p, Z, cell, precision)  # eager work between replays
        assert abs(eg - ee) <= 1e-13 * abs(ee), step
        assert np.max(np.abs(fg - fe)) <= 1e-12 * np.max(np.abs(fe)) + 1e-15, step
        assert np.max(np.abs(sg - se)) <= 1e-12 * np.max(np.abs(se)) + 1e-15, step
    # a different cell and atom count re-capture
    p2, Z2, c2 = S.si_diamond(2, jitter=0.02, seed=9)
    hb = d3_lib.Handle(np.unique(np.concatenate([Z, Z2])), len(Z) + len(Z2), precision=precision, **KW)
    hb.set_graph(True)
    for _ in range(3):
        for (pp, zz, cc) in ((pos, Z, cell), (p2, Z2, c2)):
            hb.set_atoms(pp, zz)
            hb.set_cell(cc)
            e, f, st = hb.compute()
            ee, fe, se = eager(pp, zz, cc, precision)
            assert abs(e - ee) <= 1e-13 * abs(ee)
            assert np.max(np.abs(f - fe)) <= 1e-12 * np.max(np.abs(fe)) + 1e-15
    hg.close()
    hb.close()


def test_graph_batch_replays():
    counts, pos, Z, cells = S.triclinic_batch(12, seed=4, n_max=400)
    n = int(counts.sum())
    h = d3_lib.Handle(np.unique(Z), n, **KW)
    e0, f0, s0 = h.compute_batch(counts, pos, Z, cells)
    h.set_graph(True)
    for _ in range(4):
        e, f, st = h.compute_batch(counts, pos, Z, cells)
        assert np.array_equal(e, e0)
        assert np.max(np.abs(f - f0)) <= 1e-12 * np.max(np.abs(f0))
        assert np.max(np.abs(st - s0)) <= 1e-12 * np.max(np.abs(s0))
    h.close()
