"""Per-pair C6 interpolation for parameter tables whose CN references do not factorise per
element.  The reference interpolates every pair over the pair's own 25 (or 49) entries of
c6ab[A][B][a][b] (src/d3_kernel.cu:709-800); the engine's fast path factorises that sum into
per-atom Gaussian weights, which holds when every element has one set of CN references and the
valid entries of a pair are all combinations of its elements' references (true of the D3 tables).
A table that breaks either property switches a handle to the per-pair form; these tests build
such tables (a CN reference that depends on the partner element, and an entry missing from an
otherwise complete pair block, both kept symmetric under A <-> B) and check the GPU against the
FP64 oracle, whose interpolation is the reference's literal per-pair form.  The per-pair form is
also forced on a factorising table and checked against the factorised path."""

import numpy as np
import pytest

import d3_lib
import d3_oracle
import d3_structures as S

KW = dict(cutoff=60.0, cn_cutoff=30.0, atm_cutoff=20.0, functional=0)


def _cells():
    pos, Z, cell = S.fcc_cu(2, jitter=0.05, seed=4)
    Z = Z.copy()
    Z[::4] = 1
    yield "Cu24H8", pos, Z, cell
    rng = np.random.Generator(np.random.PCG64(21))
    cell = np.array([[7.6, 0.0, 0.0], [0.9, 7.9, 0.0], [-0.6, 0.4, 8.3]])
    pos, Z = S.random_packing(18, cell, rng, elements=(1, 8, 22), d_min=1.4)
    yield "TiOH18", pos, np.asarray(Z, np.uint16), cell


CELLS = list(_cells())


def _non_factorising(t0):
    """The table with (1) Cu's and Ti's CN references shifted by 0.3 in their pairs with H only
    and (2) the entries (Cu ref 2, H ref 4), (Ti ref 3, O ref 1) removed (CN references -1),
    each change mirrored in the (B, A) block."""
    c6 = t0["c6ab"].copy()
    zo = t0["zoff"]
    H, O, Ti, Cu = 1 - zo, 8 - zo, 22 - zo, 29 - zo
    for A in (Cu, Ti):
        for a in range(t0["nref"]):
            for b in range(t0["nref"]):
                if c6[A, H, a, b, 1] > -1 and a > 0:
                    c6[A, H, a, b, 1] += 0.3
                    c6[H, A, b, a, 2] += 0.3
    for A, B, a, b in ((Cu, H, 2, 4), (Ti, O, 3, 1)):
        assert c6[A, B, a, b, 1] > -1 and c6[A, B, a, b, 2] > -1
        c6[A, B, a, b, 1:] = -1.0
        c6[B, A, b, a, 1:] = -1.0
    return c6


@pytest.fixture
def nonfact_table(tmp_path):
    import pack_table
    t0 = d3_lib.table()
    path = str(tmp_path / "nonfact.d3tab")
    pack_table.write_table(path, _non_factorising(t0), t0["r0ab"], t0["rcov"], t0["r2r4"], zoff=t0["zoff"],
                           zmax=t0["zmax"])
    d3_lib.load_table(path)
    try:
        yield d3_lib.table()
    finally:
        d3_lib.load_table(None)


def test_modified_table_does_not_factorise(table):
    """CPU: the modified table breaks both factorisation properties for the test cells' pairs."""
    c6 = _non_factorising(table)
    zo = table["zoff"]
    H, Cu = 1 - zo, 29 - zo
    # partner-dependent CN reference of Cu
    assert c6[Cu, H, 1, 0, 1] != c6[Cu, Cu, 1, 0, 1]
    # a missing entry whose row and column references are valid elsewhere in the block
    assert c6[Cu, H, 2, 4, 1] == -1 and c6[Cu, H, 2, 3, 1] > -1 and c6[Cu, H, 1, 4, 1] > -1
    # symmetric under (A, a) <-> (B, b)
    np.testing.assert_array_equal(c6[..., 0], np.swapaxes(np.swapaxes(c6[..., 0], 0, 1), 2, 3))
    np.testing.assert_array_equal(c6[..., 1], np.swapaxes(np.swapaxes(c6[..., 2], 0, 1), 2, 3))


def _run(pos, Z, cell, prec, damping, env_pairwise=False, monkeypatch=None):
    if env_pairwise:
        monkeypatch.setenv("D3_C6_PAIRWISE", "1")
    try:
        h = d3_lib.Handle(np.unique(Z), len(Z), precision=prec, damping=damping, **KW)
    finally:
        if env_pairwise:
            monkeypatch.delenv("D3_C6_PAIRWISE")
    try:
        h.set_atoms(pos, Z)
        h.set_cell(cell)
        e, f, s = h.compute()
        cn, g, ea = h.atom_data()
        return dict(pairwise=h.c6_pairwise(), e=e, f=f, s=s, cn=cn, g=g, variant=h.atm_variant())
    finally:
        h.close()


@pytest.mark.gpu
@pytest.mark.parametrize("damping", [0, 1], ids=["zero", "bj"])
@pytest.mark.parametrize("case", CELLS, ids=[c[0] for c in CELLS])
def test_non_factorising_table_parity(case, damping, nonfact_table):
    name, pos, Z, cell = case
    print(f"case pairwise {name} damping={damping} atoms={len(Z)}", flush=True)
    o = d3_oracle.compute(pos, Z, cell, nonfact_table, damping=damping, **KW)
    r = _run(pos, Z, cell, d3_lib.D3X_FP64, damping)
    assert r["pairwise"] and not (r["variant"] & 2)
    assert abs(r["e"] - o["energy"]) <= 1e-12 * abs(o["energy"])
    assert np.max(np.abs(r["f"] - o["forces"])) <= 1e-9
    assert np.max(np.abs(r["s"] - o["stress"])) <= 1e-12 * np.max(np.abs(o["stress"])) + 1e-12
    np.testing.assert_allclose(r["cn"], o["cn"], rtol=0, atol=1e-11)
    np.testing.assert_allclose(r["g"], o["dedcn"], rtol=0, atol=1e-10 * np.max(np.abs(o["dedcn"])) + 1e-20)
    r32 = _run(pos, Z, cell, d3_lib.D3X_FP32, damping)
    assert r32["pairwise"]
    assert abs(r32["e"] - o["energy"]) <= 2e-6 * abs(o["energy"])
    fs = np.max(np.abs(o["forces"]))
    assert np.max(np.abs(r32["f"] - o["forces"])) <= 2e-5 * fs + 1e-7


@pytest.mark.gpu
def test_non_factorising_table_changes_the_result(nonfact_table):
    """The modifications reach the energy: the factorising table gives a different one (so the
    parity above is not met by ignoring them)."""
    name, pos, Z, cell = CELLS[0]
    r = _run(pos, Z, cell, d3_lib.D3X_FP64, 1)
    d3_lib.load_table(None)
    r0 = _run(pos, Z, cell, d3_lib.D3X_FP64, 1)
    assert r["pairwise"] and not r0["pairwise"]
    assert abs(r["e"] - r0["e"]) > 1e-6 * abs(r0["e"])


@pytest.mark.gpu
@pytest.mark.parametrize("case", CELLS, ids=[c[0] for c in CELLS])
def test_forced_pairwise_equals_factorised(case, monkeypatch):
    """D3_C6_PAIRWISE=1 on a factorising table: in FP64 the per-pair form equals the factorised
    one to rounding (1e-12 energy, 1e-10 of the largest force); in FP32 it meets the FP32 bar
    against the oracle (2e-6 energy, 2e-5 of the largest force + 1e-7 eV/A)."""
    name, pos, Z, cell = case
    print(f"case forced-pairwise {name}", flush=True)
    o = d3_oracle.compute(pos, Z, cell, d3_lib.table(), damping=1, **KW)
    fs = np.max(np.abs(o["forces"]))
    a = _run(pos, Z, cell, d3_lib.D3X_FP64, 1)
    b = _run(pos, Z, cell, d3_lib.D3X_FP64, 1, env_pairwise=True, monkeypatch=monkeypatch)
    assert b["pairwise"] and not a["pairwise"]
    assert abs(a["e"] - b["e"]) <= 1e-12 * abs(a["e"])
    assert np.max(np.abs(a["f"] - b["f"])) <= 1e-10 * fs
    assert abs(b["e"] - o["energy"]) <= 1e-12 * abs(o["energy"])
    assert np.max(np.abs(b["f"] - o["forces"])) <= 1e-9
    b32 = _run(pos, Z, cell, d3_lib.D3X_FP32, 1, env_pairwise=True, monkeypatch=monkeypatch)
    assert b32["pairwise"]
    assert abs(b32["e"] - o["energy"]) <= 2e-6 * abs(o["energy"])
    assert np.max(np.abs(b32["f"] - o["forces"])) <= 2e-5 * fs + 1e-7


@pytest.mark.gpu
def test_three_body_network_needs_factorised_c6(nonfact_table):
    name, pos, Z, cell = CELLS[0]
    h = d3_lib.Handle(np.unique(Z), len(Z), precision=d3_lib.D3X_FP32, damping=1, **KW)
    try:
        assert h.c6_pairwise()
        with pytest.raises(d3_lib.D3Error):
            h.set_atm_mode(2)  # exact three-body plus features
    finally:
        h.close()
