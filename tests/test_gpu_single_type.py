"""Kernels specialised for one element type against the general kernels: the two-body kernel
with the centre's C6 row and damping constants in registers, the CN and chain-rule kernels
without per-pair type gathers, and the three-body kernel with constant type indices.  The same
single-element cell runs through a handle made for its element alone (specialised kernels) and
through a handle made for that element and hydrogen, which the cell does not contain (general
kernels): FP64 results agree to 1e-13 (energy, stress) and 1e-12 of the largest force, FP32
results to 1e-7 and 1e-6 (the two kernels round the same terms in a different order)."""

import numpy as np
import pytest

import d3_lib
import d3_structures as S

pytestmark = pytest.mark.gpu


def cells():
    si = S.si_diamond(nrep=2, jitter=0.02, seed=5)
    cu = S.fcc_cu(nrep=3, jitter=0.05, seed=6)
    return {"si64": si, "cu108": cu}


@pytest.mark.parametrize("name", ["si64", "cu108"])
@pytest.mark.parametrize("precision", [d3_lib.D3X_FP64, d3_lib.D3X_FP32])
def test_single_type_kernels_equal_general(name, precision):
    pos, Z, cell = cells()[name]
    print(f"case single-type {name} precision={precision}", flush=True)
    kw = dict(cutoff=60.0, cn_cutoff=30.0, atm_cutoff=25.0, damping=1, functional=0, precision=precision)
    out = []
    for els in ([int(Z[0])], [1, int(Z[0])]):
        h = d3_lib.Handle(els, len(Z), **kw)
        h.set_atoms(pos, Z)
        h.set_cell(cell)
        e, f, st = h.compute()
        cn, g, _ = h.atom_data()
        h.close()
        out.append((e, f, st, cn, g))
    (e1, f1, s1, cn1, g1), (e2, f2, s2, cn2, g2) = out
    te, tf = (1e-13, 1e-12) if precision == d3_lib.D3X_FP64 else (1e-7, 1e-6)
    assert abs(e1 - e2) <= te * abs(e2)
    assert np.max(np.abs(f1 - f2)) <= tf * np.max(np.abs(f2))
    assert np.max(np.abs(s1 - s2)) <= te * 10 * np.max(np.abs(s2))
    np.testing.assert_allclose(cn1, cn2, rtol=te * 10, atol=0)
    np.testing.assert_allclose(g1, g2, rtol=tf, atol=tf * np.max(np.abs(g2)))
