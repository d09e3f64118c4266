"""GPU parity: the HIP path through the C ABI against the FP64 oracle on the same inputs.

Bars: plain FP64 path within 1e-12 relative energy and 1e-9
eV/Angstrom forces of the oracle; stress within 1e-12 relative to its largest component (plus
1e-12 eV/A^3 absolute); CN and dE/dCN to 1e-10 relative.  The FP32 path is held to 2e-6 relative
energy and 2e-5 of the largest force component; its measured error is printed per system.
"""

import numpy as np
import pytest

import d3_lib
import d3_oracle
import systems

pytestmark = pytest.mark.gpu

SYSTEMS = systems.parity_systems()


def _oracle(s, table):
    return d3_oracle.compute(s["pos"], s["Z"], s["cell"], table, cutoff=s["cutoff"], cn_cutoff=s["cn_cutoff"],
                             atm_cutoff=s.get("atm", 0.0), damping=s["damping"], functional=s["functional"])


def _gpu(s, precision):
    h = d3_lib.Handle(s["Z"], len(s["Z"]), cutoff=s["cutoff"], cn_cutoff=s["cn_cutoff"],
                      atm_cutoff=s.get("atm", 0.0), damping=s["damping"], functional=s["functional"],
                      precision=precision)
    h.set_atoms(s["pos"], s["Z"])
    h.set_cell(s["cell"])
    e, f, st = h.compute()
    cn, g, ea = h.atom_data()
    h.close()
    return dict(energy=e, forces=f, stress=st, cn=cn, dedcn=g, atom_energy=ea)


_cache = {}


def oracle_result(s, table):
    if s["name"] not in _cache:
        _cache[s["name"]] = _oracle(s, table)
    return _cache[s["name"]]


@pytest.mark.parametrize("s", SYSTEMS, ids=[s["name"] for s in SYSTEMS])
def test_fp64_matches_oracle(s, table):
    print(f"case {s['name']} n={len(s['Z'])} fp64", flush=True)
    o = oracle_result(s, table)
    g = _gpu(s, d3_lib.D3X_FP64)
    assert abs(g["energy"] - o["energy"]) <= 1e-12 * abs(o["energy"]), (g["energy"], o["energy"])
    assert np.max(np.abs(g["forces"] - o["forces"])) <= 1e-9
    smax = np.max(np.abs(o["stress"]))
    assert np.max(np.abs(g["stress"] - o["stress"])) <= 1e-12 * smax + 1e-12
    # per-atom values: relative to each array's largest magnitude, with a floor far below any
    # physical size (values near zero carry the rounding noise of the oracle's Z/W
    # cancellation; dE/dCN vanishes when one reference dominates the weights)
    for k, floor in (("cn", 1e-14), ("dedcn", 1e-20), ("atom_energy", 1e-15)):
        scale = np.max(np.abs(o[k]))
        assert np.max(np.abs(g[k] - o[k])) <= 1e-11 * scale + floor, k


@pytest.mark.parametrize("s", SYSTEMS, ids=[s["name"] for s in SYSTEMS])
def test_fp32_error_vs_oracle(s, table):
    print(f"case {s['name']} n={len(s['Z'])} fp32", flush=True)
    o = oracle_result(s, table)
    g = _gpu(s, d3_lib.D3X_FP32)
    rel_e = abs(g["energy"] - o["energy"]) / abs(o["energy"])
    fmax = np.max(np.abs(o["forces"]))
    df = np.max(np.abs(g["forces"] - o["forces"]))
    smax = np.max(np.abs(o["stress"]))
    ds = np.max(np.abs(g["stress"] - o["stress"]))
    print(f"  fp32 rel dE {rel_e:.2e}  max dF {df:.2e} eV/A (|F|max {fmax:.2e})  max dS {ds:.2e} (|S|max {smax:.2e})")
    assert rel_e <= 2e-6
    assert df <= 2e-5 * fmax + 1e-7
    assert ds <= 2e-5 * smax + 1e-9


def test_batch_equals_single(table):
    """One batched evaluation of several cells equals the cells evaluated one by one."""
    cases = [systems.si64(seed=11), systems.triclinic_mixed(seed=12), systems.water_small(seed=13)]
    # common handle parameters: union of elements, one cutoff set
    els = np.unique(np.concatenate([c["Z"] for c in cases]))
    kw = dict(cutoff=60.0, cn_cutoff=30.0, atm_cutoff=25.0, damping=1, functional=0)
    total = sum(len(c["Z"]) for c in cases)
    hb = d3_lib.Handle(els, total, **kw)
    eb, fb, sb = hb.compute_batch([len(c["Z"]) for c in cases], np.concatenate([c["pos"] for c in cases]),
                                  np.concatenate([c["Z"] for c in cases]), np.stack([c["cell"] for c in cases]))
    off = 0
    for k, c in enumerate(cases):
        hs = d3_lib.Handle(els, len(c["Z"]), **kw)
        hs.set_atoms(c["pos"], c["Z"])
        hs.set_cell(c["cell"])
        e, f, st = hs.compute()
        n = len(c["Z"])
        assert abs(e - eb[k]) <= 1e-13 * abs(e)
        assert np.max(np.abs(f - fb[off:off + n])) <= 1e-12
        assert np.max(np.abs(st - sb[k])) <= 1e-13 * np.max(np.abs(st)) + 1e-15
        off += n


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_drop_in_float_api(table, precision, monkeypatch):
    """The reference-compatible float API (d3_cffi.D3Calculator): by default the FP32 path (within
    its 2e-6 / 2e-5 bar of the oracle), with D3_PRECISION=fp64 the FP64 results rounded to float."""
    import d3_cffi
    if precision == "fp64":
        monkeypatch.setenv("D3_PRECISION", "fp64")
    else:
        monkeypatch.delenv("D3_PRECISION", raising=False)
    s = systems.SMALL
    calc = d3_cffi.D3Calculator(list(s["Z"]), len(s["Z"]), cutoff_radius=40.0, cn_cutoff_radius=40.0,
                                damping_type=0, functional_type=1)
    calc.set_atoms(s["pos"].astype(np.float32), list(s["Z"]))
    calc.set_cell(s["cell"].astype(np.float32))
    e, f, st = calc.compute()
    calc.clear()
    o = oracle_result(s, table)
    # float32 rounding of Angstrom inputs (the test coordinates are float32 values already)
    te, tf = (1e-6, 1e-6) if precision == "fp64" else (2e-6, 2e-5)
    assert abs(e - o["energy"]) <= te * abs(o["energy"])
    assert np.max(np.abs(f - o["forces"])) <= tf * np.max(np.abs(o["forces"])) + 1e-9
    assert f.dtype == np.float32 and f.shape == (10, 3) and st.shape == (3, 3)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_drop_in_md_loop(table, precision, monkeypatch):
    """The reference's MD pattern (src/test_d3_cffi.py:79-123): one D3Calculator with the
    three-body term, then per step set_atoms -> set_cell -> compute -> clear on coordinates
    displaced by U(-0.01, 0.01) A per component and a slowly strained cell; every step against the
    oracle on the same float inputs (FP64 at the float output's resolution, FP32 at its bar)."""
    import d3_cffi
    if precision == "fp64":
        monkeypatch.setenv("D3_PRECISION", "fp64")
    else:
        monkeypatch.delenv("D3_PRECISION", raising=False)
    s = systems.triclinic_mixed(n=36, seed=41)
    Z = s["Z"]
    kw = dict(cutoff=50.0, cn_cutoff=30.0, atm_cutoff=25.0, damping=1, functional=0)
    print(f"case md-loop {precision} atoms={len(Z)}", flush=True)
    calc = d3_cffi.D3Calculator(list(Z), len(Z), cutoff_radius=kw["cutoff"], cn_cutoff_radius=kw["cn_cutoff"],
                                damping_type=1, functional_type=0, atm_cutoff=kw["atm_cutoff"])
    rng = np.random.Generator(np.random.PCG64(7))
    pos = s["pos"].copy()
    te, tf = (1e-6, 1e-6) if precision == "fp64" else (2e-6, 2e-5)
    for step in range(10):
        pos = pos + rng.uniform(-0.01, 0.01, pos.shape)
        cell = s["cell"] @ (np.eye(3) + 1e-3 * step * np.array([[1, 0.2, 0], [0, -0.5, 0], [0.1, 0, 0.3]]))
        p32, c32 = pos.astype(np.float32), cell.astype(np.float32)
        calc.set_atoms(p32, list(Z))
        calc.set_cell(c32)
        e, f, st = calc.compute()
        calc.clear()
        o = d3_oracle.compute(p32.astype(np.float64), Z, c32.astype(np.float64), table, **kw)
        assert abs(e - o["energy"]) <= te * abs(o["energy"]), step
        assert np.max(np.abs(f - o["forces"])) <= tf * np.max(np.abs(o["forces"])) + 1e-9, step
        assert np.max(np.abs(st - o["stress"])) <= tf * np.max(np.abs(o["stress"])) + 1e-9, step
    del calc


@pytest.mark.parametrize("precision", [d3_lib.D3X_FP64, d3_lib.D3X_FP32])
def test_smooth_cn_cutoff(table, precision):
    """The smooth CN cutoff option against the oracle with the same option, three-body term on, on a
    CN cutoff short enough (7 bohr) that the smooth factor changes CN and forces."""
    s = systems.triclinic_mixed(n=40, seed=71)
    kw = dict(cutoff=40.0, cn_cutoff=7.0, atm_cutoff=25.0, damping=1, functional=0)
    print(f"case smooth CN precision={precision}", flush=True)
    h = d3_lib.Handle(np.unique(s["Z"]), len(s["Z"]), precision=precision, **kw)
    h.set_cn_smooth(True)
    h.set_atoms(s["pos"], s["Z"])
    h.set_cell(s["cell"])
    e, f, st = h.compute()
    cn, g, _ = h.atom_data()
    h.close()
    o = d3_oracle.compute(s["pos"], s["Z"], s["cell"], table, smooth_cn=True, **kw)
    sharp = d3_oracle.compute(s["pos"], s["Z"], s["cell"], table, **kw)
    assert np.max(np.abs(sharp["cn"] - o["cn"])) > 1e-3  # the option matters in this cell
    if precision == d3_lib.D3X_FP64:
        assert abs(e - o["energy"]) <= 1e-12 * abs(o["energy"])
        assert np.max(np.abs(f - o["forces"])) <= 1e-9
        assert np.max(np.abs(st - o["stress"])) <= 1e-12 * np.max(np.abs(o["stress"])) + 1e-12
        np.testing.assert_allclose(cn, o["cn"], rtol=1e-12, atol=1e-14)
    else:
        assert abs(e - o["energy"]) <= 2e-6 * abs(o["energy"])
        assert np.max(np.abs(f - o["forces"])) <= 2e-5 * np.max(np.abs(o["forces"])) + 1e-7


@pytest.mark.parametrize("atm", [0.0, 40.0])
def test_repeat_is_stable(table, atm):
    """ResultStable (test/d3_test.cpp:89-129, 1e-7 eV / 1e-5 there): without the three-body term
    repeated evaluations agree bit for bit; with it the energy still does, while forces and stress
    carry the FP64-atomic scatter of the triangle terms (order-dependent at the 1e-16 level)."""
    s = systems.si64(seed=21)
    h = d3_lib.Handle(s["Z"], len(s["Z"]), cutoff=s["cutoff"], cn_cutoff=s["cn_cutoff"], atm_cutoff=atm,
                      damping=1, functional=0)
    h.set_atoms(s["pos"], s["Z"])
    h.set_cell(s["cell"])
    e0, f0, s0 = h.compute()
    for _ in range(8):
        e, f, st = h.compute()
        assert e == e0
        if atm == 0.0:
            assert np.array_equal(f, f0) and np.array_equal(st, s0)
        else:
            assert np.max(np.abs(f - f0)) <= 1e-13 * np.max(np.abs(f0))
            assert np.max(np.abs(st - s0)) <= 1e-13 * np.max(np.abs(s0))
