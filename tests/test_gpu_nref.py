"""Reference counts: a handle keeps only the C6 references its elements have (trailing references
with no valid entry for any of its elements carry zero Gaussian weight), and the three-body kernel
is instantiated per count.  Cells whose element sets give 1, 3 and 5 references (the 95 x 5 table)
and 1 to 7 references (the 103 x 7 extended table, slots Z - 1 as the reference's IDX_ELEM with
USE_EXTENDED_PARAMETERS, src/d3_buffer.cu:11-19, src/constants_include.h:25-30) are checked
against the FP64 oracle with the three-body term on (FP64 at the parity bar, FP32 at 2e-6
relative energy); the CPU tests pin which counts the cells cover."""

import numpy as np
import pytest

import d3_lib
import d3_oracle
import d3_structures as S


def _cells():
    rng = np.random.Generator(np.random.PCG64(11))
    cell = np.eye(3) * 7.5
    pos, _ = S.random_packing(14, cell, rng, d_min=1.3)
    yield "C14", pos, np.full(len(pos), 6, np.uint16), cell, 22.0
    pos, Z, cell = S.fcc_cu(2, jitter=0.05, seed=4)
    yield "Cu32", pos, Z, cell, 20.0
    Z = Z.copy()
    Z[::4] = 1
    yield "Cu24H8", pos, Z, cell, 20.0


CELLS = list(_cells())


def _extended_cells():
    """Element sets of the extended table whose reference counts are 1, 2, ..., 7 (B, C, Ga, Cu,
    Cu + N, Cu + S, Cu + O + H)."""
    rng = np.random.Generator(np.random.PCG64(12))
    cell = np.eye(3) * 7.5
    for name, z in (("B14", 5), ("C14", 6), ("Ga14", 31)):
        pos, _ = S.random_packing(14, cell, rng, d_min=1.3)
        yield "ext-" + name, pos, np.full(len(pos), z, np.uint16), cell, 22.0
    pos, Z, cucell = S.fcc_cu(2, jitter=0.05, seed=5)
    yield "ext-Cu32", pos, Z, cucell, 20.0
    for name, sub in (("Cu24N8", (7, 7)), ("Cu24S8", (16, 16)), ("Cu24O4H4", (8, 1))):
        Zm = Z.copy()
        Zm[0::8] = sub[0]
        Zm[4::8] = sub[1]
        yield "ext-" + name, pos, Zm, cucell, 20.0


EXT_CELLS = list(_extended_cells())


@pytest.fixture
def extended_table():
    """The 103 x 7 synthetic table, active for the test only (handles copy it at init)."""
    d3_lib.load_table("synthetic-extended")
    try:
        yield d3_lib.table()
    finally:
        d3_lib.load_table(None)


def effective_nref(table, Z):
    """Largest (last valid reference + 1) over the elements of Z: a reference a of element A is
    valid when C6ref(A, A, a, a) carries a CN reference (> -1)."""
    c6 = table["c6ab"]
    n = 1
    for z in np.unique(Z):
        s = int(z) - table["zoff"]
        for a in range(table["nref"]):
            if c6[s, s, a, a, 1] > -1:
                n = max(n, a + 1)
    return n


def test_cells_cover_reference_counts(table):
    counts = [effective_nref(table, Z) for _, _, Z, _, _ in CELLS]
    assert counts[0] < counts[1] < counts[2] <= table["nref"], counts
    assert counts[0] == 1 and counts[2] == table["nref"]


def test_extended_cells_cover_every_reference_count(extended_table):
    counts = [effective_nref(extended_table, Z) for _, _, Z, _, _ in EXT_CELLS]
    assert counts == list(range(1, extended_table["nref"] + 1)), counts


@pytest.mark.gpu
@pytest.mark.parametrize("case", EXT_CELLS, ids=[c[0] for c in EXT_CELLS])
def test_extended_table_parity(case, extended_table):
    _parity(case, extended_table)


@pytest.mark.gpu
@pytest.mark.parametrize("case", CELLS, ids=[c[0] for c in CELLS])
def test_reference_count_parity(case, table):
    _parity(case, table)


def _parity(case, table):
    name, pos, Z, cell, rc3 = case
    print(f"case nref {name} table={table['name']} atoms={len(Z)} nref={effective_nref(table, Z)}", flush=True)
    kw = dict(cutoff=60.0, cn_cutoff=30.0, atm_cutoff=rc3, damping=1, functional=0)
    o = d3_oracle.compute(pos, Z, cell, table, **kw)
    out = {}
    for prec in (d3_lib.D3X_FP64, d3_lib.D3X_FP32):
        h = d3_lib.Handle(np.unique(Z), len(Z), precision=prec, **kw)
        h.set_atoms(pos, Z)
        h.set_cell(cell)
        out[prec] = h.compute()
        cn, g, ea = h.atom_data()
        h.close()
        if prec == d3_lib.D3X_FP64:
            np.testing.assert_allclose(g, o["dedcn"], rtol=0, atol=1e-10 * np.max(np.abs(o["dedcn"])) + 1e-20)
    e, f, s = out[d3_lib.D3X_FP64]
    assert abs(e - o["energy"]) <= 1e-12 * abs(o["energy"])
    assert np.max(np.abs(f - o["forces"])) <= 1e-9
    assert np.max(np.abs(s - o["stress"])) <= 1e-12 * np.max(np.abs(o["stress"])) + 1e-12
    e32 = out[d3_lib.D3X_FP32][0]
    assert abs(e32 - o["energy"]) <= 2e-6 * abs(o["energy"])
