"""Neighbour state on the GPU: the three-body forward lists give the stencil walk's results (the
same neighbours in the same order, so the same sums), and with a skin the handle keeps its bin
order and lists across set_atoms calls while the atoms stay within half the skin -- along an
MD-like trajectory every step equals a handle that rebuilds everything (FP64 at the parity bar),
an unchanged input gives the rebuild's bits, and moving an atom beyond half the skin, changing the
cell or evaluating a batch rebuilds."""

import ctypes as C

import numpy as np
import pytest

import d3_lib
import d3_oracle
import d3_structures as S
import systems

pytestmark = pytest.mark.gpu

KW = dict(cutoff=50.0, cn_cutoff=30.0, atm_cutoff=30.0, damping=1, functional=0)


def _parity(out, ref, rel_e=1e-12, df=1e-9):
    e, f, s = out
    er, fr, sr = ref
    assert abs(e - er) <= rel_e * abs(er)
    assert np.max(np.abs(f - fr)) <= df
    assert np.max(np.abs(s - sr)) <= 1e-12 * np.max(np.abs(sr)) + 1e-14


@pytest.mark.parametrize("precision", [d3_lib.D3X_FP64, d3_lib.D3X_FP32])
def test_forward_lists_equal_stencil_walk(precision):
    """Si 512 (each centre split into several three-body work items) with and without lists."""
    pos, Z, cell = S.si_diamond(4, jitter=0.02, seed=3)
    print(f"case lists vs walk si512 precision={precision}", flush=True)
    out = []
    for mode in (0, 1):
        h = d3_lib.Handle(Z, len(Z), cutoff=float(np.sqrt(9000.0)), cn_cutoff=40.0, atm_cutoff=40.0, damping=1,
                          functional=0, precision=precision)
        h.set_neighbour_lists(mode)
        h.set_atoms(pos, Z)
        h.set_cell(cell)
        out.append(h.compute() + h.atom_data())
        h.close()
    (e0, f0, s0, cn0, g0, ea0), (e1, f1, s1, cn1, g1, ea1) = out
    # the total energy (anchored triangles, fixed order) and CN are equal bit for bit; per-atom
    # energies, forces, dE/dCN and stress carry the three-body term's FP64 atomic scatter
    # (order-dependent at the last bits)
    assert e0 == e1 and np.array_equal(cn0, cn1)
    assert np.max(np.abs(ea1 - ea0)) <= 1e-13 * np.max(np.abs(ea0))
    assert np.max(np.abs(f1 - f0)) <= 1e-13 * np.max(np.abs(f0))
    assert np.max(np.abs(g1 - g0)) <= 1e-13 * np.max(np.abs(g0))
    assert np.max(np.abs(s1 - s0)) <= 1e-13 * np.max(np.abs(s0))


def test_skin_reuse_along_a_trajectory(table):
    s = systems.triclinic_mixed(n=40, seed=61)
    pos0, Z, cell = s["pos"], s["Z"], s["cell"]
    skin = 1.0  # bohr: rebuild once an atom has moved more than 0.5 bohr (0.26 A)
    print(f"case skin reuse trajectory atoms={len(Z)}", flush=True)
    h = d3_lib.Handle(np.unique(Z), len(Z), **KW)
    h.set_skin(skin)
    ref = d3_lib.Handle(np.unique(Z), len(Z), **KW)
    rng = np.random.Generator(np.random.PCG64(5))
    pos = pos0.copy()
    for step in range(16):
        pos = pos + rng.uniform(-0.05, 0.05, pos.shape)
        h.set_atoms(pos, Z)
        h.set_cell(cell)
        out = h.compute()
        ref.set_atoms(pos, Z)
        ref.set_cell(cell)
        _parity(out, ref.compute())
    rebuilds, reuses = h.neighbour_stats()
    print(f"  rebuilds {rebuilds}, reuses {reuses}", flush=True)
    assert rebuilds >= 2 and reuses >= 4 and rebuilds + reuses == 16
    # and against the oracle at the last step
    o = d3_oracle.compute(pos, Z, cell, table, **KW)
    _parity(out, (o["energy"], o["forces"], o["stress"]))
    h.close()
    ref.close()


def test_skin_reuse_of_an_unchanged_input_is_bitwise():
    s = systems.triclinic_mixed(n=32, seed=62)
    pos, Z, cell = s["pos"], s["Z"], s["cell"]
    print("case skin reuse unchanged input", flush=True)
    kw = dict(KW, atm_cutoff=0.0)  # without the three-body scatter forces are fixed-order sums too
    for atm in (0.0, 30.0):
        h = d3_lib.Handle(np.unique(Z), len(Z), **dict(kw, atm_cutoff=atm))
        h.set_skin(0.8)
        h.set_atoms(pos, Z)
        h.set_cell(cell)
        e0, f0, s0 = h.compute()
        h.set_atoms(pos.copy(), Z)  # same positions: the kept state serves
        h.set_cell(cell.copy())
        e1, f1, s1 = h.compute()
        assert h.neighbour_stats() == (1, 1)
        assert e1 == e0
        if atm == 0.0:
            assert np.array_equal(f1, f0) and np.array_equal(s1, s0)
        else:
            assert np.max(np.abs(f1 - f0)) <= 1e-13 * np.max(np.abs(f0))
        h.close()


def test_what_forces_a_rebuild():
    s = systems.triclinic_mixed(n=24, seed=63)
    pos, Z, cell = s["pos"], s["Z"], s["cell"]
    print("case skin rebuild triggers", flush=True)
    h = d3_lib.Handle(np.unique(Z), 2 * len(Z), **KW)
    h.set_skin(1.0)
    h.set_atoms(pos, Z)
    h.set_cell(cell)
    h.compute()  # rebuild 1
    p = pos.copy()
    p[3, 0] += 0.5 * 1.0 * 0.52917726 * 0.9  # within half the skin
    h.set_atoms(p, Z)
    h.compute()  # reuse
    assert h.neighbour_stats() == (1, 1)
    p[3, 0] += 0.5 * 0.52917726 * 0.2  # now beyond half the skin from the rebuild
    h.set_atoms(p, Z)
    h.compute()  # rebuild 2
    assert h.neighbour_stats() == (2, 1)
    h.set_cell(cell * 1.001)  # another cell
    h.set_atoms(p, Z)
    h.compute()  # rebuild 3
    assert h.neighbour_stats() == (3, 1)
    h.compute_batch([len(Z), len(Z)], np.concatenate([p, pos]), np.concatenate([Z, Z]), np.stack([cell, cell]))
    assert h.neighbour_stats() == (4, 1)  # batches rebuild
    h.close()


def test_skin_from_environment_in_the_drop_in(monkeypatch, table):
    """D3_SKIN=<bohr> gives D3Calculator handles the kept neighbour state."""
    import d3_cffi
    monkeypatch.setenv("D3_SKIN", "1.0")
    monkeypatch.setenv("D3_PRECISION", "fp64")
    s = systems.triclinic_mixed(n=30, seed=64)
    Z = s["Z"]
    print("case skin drop-in", flush=True)
    calc = d3_cffi.D3Calculator(list(Z), len(Z), cutoff_radius=KW["cutoff"], cn_cutoff_radius=KW["cn_cutoff"],
                                damping_type=1, functional_type=0, atm_cutoff=KW["atm_cutoff"])
    rng = np.random.Generator(np.random.PCG64(9))
    pos = s["pos"].copy()
    for step in range(6):
        pos = pos + rng.uniform(-0.01, 0.01, pos.shape)
        p32, c32 = pos.astype(np.float32), s["cell"].astype(np.float32)
        calc.set_atoms(p32, list(Z))
        calc.set_cell(c32)
        e, f, st = calc.compute()
        calc.clear()
        o = d3_oracle.compute(p32.astype(np.float64), Z, c32.astype(np.float64), table, **KW)
        assert abs(e - o["energy"]) <= 1e-6 * abs(o["energy"])
        assert np.max(np.abs(f - o["forces"])) <= 1e-6 * np.max(np.abs(o["forces"])) + 1e-9
    a, b = C.c_int(), C.c_int()
    h = calc.handle if isinstance(calc.handle, int) else C.cast(calc.handle, C.c_void_p).value
    assert d3_lib.load().d3x_neighbour_stats(C.c_void_p(h), C.byref(a), C.byref(b)) == 0
    assert b.value >= 3, (a.value, b.value)
    del calc
