"""GPU checks at the benchmark workloads' full sizes:

* the 512-atom Si diamond cell (the oracle's own CPU case) against the FP64 oracle at the parity bar;
* the 12,288-atom water box: translation and permutation invariance, sum of forces, stress
  symmetry, and the FP32 path against the FP64 path (<= 1e-6 relative energy);
* the 256-cell triclinic batch: per-cell results of the batched evaluation equal single-cell runs
  for a sample of cells, sum of forces per cell, FP32 against FP64 per cell;
* the 1,000,188-atom fcc Cu cell without jitter: per-atom energy, CN, dE/dCN and stress equal the
  4-atom cell's (oracle), forces vanish;
* the jittered 1,000,188-atom Cu cell and the 100,000-atom random H/C/N/O/Si/Cu cell that the
  benchmark times: at random centre atoms the oracle's per-centre passes (CN from scratch; energy,
  dE/dCN and force with the other atoms' CN and dE/dCN from the engine) against the engine's
  per-atom results, FP64 at the parity bar and FP32 at its own; the random cell also with the
  three-body network in place of the exact term (judged against the exact kernel).
"""

import numpy as np
import pytest

import d3_lib
import d3_oracle
import d3_structures as S

pytestmark = pytest.mark.gpu

RC = dict(cutoff=float(np.sqrt(9000.0)), cn_cutoff=40.0, atm_cutoff=40.0, damping=1, functional=0)


def handle(Z, n, precision=d3_lib.D3X_FP64, **kw):
    args = dict(RC)
    args.update(kw)
    return d3_lib.Handle(np.unique(Z), n, precision=precision, **args)


def run(pos, Z, cell, precision=d3_lib.D3X_FP64):
    h = handle(Z, len(Z), precision)
    h.set_atoms(pos, Z)
    h.set_cell(cell)
    out = h.compute()
    h.close()
    return out


def test_si512_matches_oracle(table):
    pos, Z, cell = S.si_diamond(4, jitter=0.02, seed=0)
    print("case si512 fp64 vs oracle", flush=True)
    e, f, st = run(pos, Z, cell)
    o = d3_oracle.compute(pos, Z, cell, table, **RC)
    assert abs(e - o["energy"]) <= 1e-12 * abs(o["energy"])
    assert np.max(np.abs(f - o["forces"])) <= 1e-9
    assert np.max(np.abs(st - o["stress"])) <= 1e-12 * np.max(np.abs(o["stress"]))


@pytest.fixture(scope="module")
def water():
    return S.water_box(4096, seed=0)


def test_water_invariances(water):
    pos, Z, cell = water
    print("case water-4096 invariances", flush=True)
    e, f, st = run(pos, Z, cell)
    fmax = np.max(np.abs(f))
    assert np.max(np.abs(f.sum(axis=0))) <= 1e-9 * fmax * len(Z)
    assert np.max(np.abs(st - st.T)) <= 1e-13 * np.max(np.abs(st))
    e2, f2, st2 = run(pos + np.array([3.3, -7.1, 12.9]), Z, cell)
    assert abs(e2 - e) <= 1e-12 * abs(e)
    assert np.max(np.abs(f2 - f)) <= 1e-9
    p = np.random.Generator(np.random.PCG64(7)).permutation(len(Z))
    e3, f3, st3 = run(pos[p], Z[p], cell)
    assert abs(e3 - e) <= 1e-12 * abs(e)
    assert np.max(np.abs(f3 - f[p])) <= 1e-9
    assert np.max(np.abs(st3 - st)) <= 1e-12 * np.max(np.abs(st))


def test_water_fp32_error(water):
    pos, Z, cell = water
    print("case water-4096 fp32 vs fp64", flush=True)
    e64, f64, s64 = run(pos, Z, cell)
    e32, f32, s32 = run(pos, Z, cell, d3_lib.D3X_FP32)
    rel = abs(e32 - e64) / abs(e64)
    print(f"  water fp32: rel dE {rel:.2e}, max dF {np.max(np.abs(f32 - f64)):.2e} eV/A, "
          f"max dS {np.max(np.abs(s32 - s64)):.2e}")
    assert rel <= 1e-6
    assert np.max(np.abs(f32 - f64)) <= 1e-5 * np.max(np.abs(f64))
    assert np.max(np.abs(s32 - s64)) <= 1e-5 * np.max(np.abs(s64))


def test_batch256_equals_single_cells():
    counts, pos, Z, cells = S.triclinic_batch(256, seed=0)
    print(f"case batch-256 atoms={len(Z)}", flush=True)
    n = int(counts.sum())
    h = handle(Z, n)
    eb, fb, sb = h.compute_batch(counts, pos, Z, cells)
    h.close()
    off = np.concatenate([[0], np.cumsum(counts)]).astype(int)
    for s in range(len(counts)):
        fs = fb[off[s]:off[s + 1]]
        assert np.max(np.abs(fs.sum(axis=0))) <= 1e-8 * max(1.0, np.max(np.abs(fs))) * counts[s]
    for s in (0, 1, 17, 100, int(np.argmax(counts)), int(np.argmin(counts))):
        sl = slice(off[s], off[s + 1])
        e, f, st = run(pos[sl], Z[sl], cells[s])
        assert abs(e - eb[s]) <= 1e-12 * abs(e)
        assert np.max(np.abs(f - fb[sl])) <= 1e-10
        assert np.max(np.abs(st - sb[s])) <= 1e-12 * np.max(np.abs(st))
    h32 = handle(Z, n, d3_lib.D3X_FP32)
    e32, _, _ = h32.compute_batch(counts, pos, Z, cells)
    h32.close()
    rel = np.max(np.abs(e32 - eb) / np.abs(eb))
    print(f"  batch fp32: max per-cell rel dE {rel:.2e}")
    assert rel <= 1e-6


def test_perfect_fcc_cu_million_atoms(table):
    """The 1,000,188-atom fcc Cu cell of the headline benchmark without its jitter: every atom is
    equivalent, so the FP64 engine's per-atom energy, CN and dE/dCN must be those of the 4-atom
    conventional cell (FP64 oracle, same cutoffs), the forces vanish and the stress (intensive)
    equals the small cell's -- a size-independent check of the full-size path, three-body term
    included (8.9e11 triangles).  The FP32 path stays within 1e-6 of the FP64 energy."""
    pos, Z, cell = S.fcc_cu(63, jitter=0.0)
    spos, sZ, scell = S.fcc_cu(1, jitter=0.0)
    print("case perfect fcc Cu: 4-atom oracle", flush=True)
    o = d3_oracle.compute(spos, sZ, scell, table, **RC)
    e_atom = o["energy"] / len(sZ)
    n = len(Z)
    print(f"case perfect fcc Cu: {n} atoms fp64", flush=True)
    h = handle(Z, n)
    h.set_atoms(pos, Z)
    h.set_cell(cell)
    e, f, st = h.compute()
    cn, g, ea = h.atom_data()
    h.close()
    assert abs(e / n - e_atom) <= 1e-12 * abs(e_atom), (e / n, e_atom)
    assert np.max(np.abs(ea - e_atom)) <= 1e-11 * abs(e_atom)
    np.testing.assert_allclose(cn, o["cn"][0], rtol=1e-12)
    np.testing.assert_allclose(g, o["dedcn"][0], rtol=1e-11, atol=1e-20)  # 0 here: one reference dominates
    assert np.max(np.abs(f)) <= 1e-9
    assert np.max(np.abs(st - o["stress"])) <= 1e-12 * np.max(np.abs(o["stress"]))
    print(f"case perfect fcc Cu: {n} atoms fp32", flush=True)
    e32 = run(pos, Z, cell, d3_lib.D3X_FP32)[0]
    assert abs(e32 - e) <= 1e-6 * abs(e)


def _centre_parity(pos, Z, cell, table, precision, n_centres, seed, kw=RC):
    """Engine per-atom results at n_centres random atoms against the oracle's per-centre passes."""
    n = len(Z)
    h = handle(Z, n, precision, **{k: v for k, v in kw.items() if k not in RC or RC[k] != v})
    h.set_atoms(pos, Z)
    h.set_cell(cell)
    e, f, st = h.compute()
    cn, g, ea = h.atom_data()
    h.close()
    cen = np.sort(np.random.Generator(np.random.PCG64(seed)).choice(n, n_centres, replace=False))
    o = d3_oracle.centers(pos, Z, cell, table, cen, cn, g, **kw)
    d = dict(cn=np.max(np.abs(cn[cen] - o["cn"]) / np.maximum(np.abs(o["cn"]), 1.0)),
             e=np.max(np.abs(ea[cen] - o["atom_energy"]) / np.abs(o["atom_energy"])),
             # dE/dCN against its own scale or, where it nearly vanishes (one reference dominating
             # every weight, as on Cu), against the per-atom energy in Hartree
             g=np.max(np.abs(g[cen] - o["dedcn"])) / max(np.max(np.abs(o["dedcn"])),
                                                         np.max(np.abs(o["atom_energy"])) / 27.211396641308),
             f=np.max(np.abs(f[cen] - o["forces"])), fmax=np.max(np.abs(o["forces"])))
    print("  max deviation: " + ", ".join(f"{k} {v:.2e}" for k, v in d.items()), flush=True)
    return d, (e, f, st)


@pytest.fixture(scope="module")
def jittered_cu():
    return S.fcc_cu(63, jitter=0.05, seed=0)  # the benchmark's input


@pytest.mark.parametrize("precision", [d3_lib.D3X_FP64, d3_lib.D3X_FP32])
def test_jittered_cu_million_atoms_at_random_centres(table, jittered_cu, precision):
    pos, Z, cell = jittered_cu
    print(f"case jittered fcc Cu {len(Z)} atoms precision={precision}: 64 centres", flush=True)
    d, _ = _centre_parity(pos, Z, cell, table, precision, 64, seed=101)
    if precision == d3_lib.D3X_FP64:
        assert d["cn"] <= 1e-12 and d["e"] <= 1e-11 and d["g"] <= 1e-10 and d["f"] <= 1e-9
    else:
        assert d["cn"] <= 2e-6 and d["e"] <= 2e-6 and d["g"] <= 2e-5 and d["f"] <= 5e-6


@pytest.fixture(scope="module")
def random100k():
    return S.random_cell(100000, seed=4242)  # the benchmark's network workload


@pytest.mark.parametrize("precision", [d3_lib.D3X_FP64, d3_lib.D3X_FP32])
def test_random_cell_100k_exact_at_random_centres(table, random100k, precision):
    pos, Z, cell = random100k
    print(f"case random cell {len(Z)} atoms precision={precision}: 96 centres", flush=True)
    d, _ = _centre_parity(pos, Z, cell, table, precision, 96, seed=102)
    if precision == d3_lib.D3X_FP64:
        assert d["cn"] <= 1e-12 and d["e"] <= 1e-11 and d["g"] <= 1e-10 and d["f"] <= 1e-9
    else:
        assert d["cn"] <= 2e-6 and d["e"] <= 2e-6 and d["g"] <= 2e-5 and d["f"] <= 2e-5 * d["fmax"] + 1e-6


def test_random_cell_100k_network_mode(random100k):
    """The benchmark's network leg as a test: the judged first evaluation either accepts the network
    (energy within 1 meV/atom and forces within the judge's RMS tolerance of the exact kernel's) or
    falls back and returns the exact kernel's results."""
    import os
    pos, Z, cell = random100k
    n = len(Z)
    weights = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "lasp-d3_amd", "surrogate",
                           "atm_surrogate.d3sur")
    print(f"case random cell {n} atoms: network mode", flush=True)
    h = handle(Z, n, d3_lib.D3X_FP32)
    h.set_atoms(pos, Z)
    h.set_cell(cell)
    e_x, f_x, s_x = h.compute()
    h.sur_load_file(weights)
    h.set_atm_mode(d3_lib.D3X_ATM_SURROGATE, 0, 1e-3)
    e_n, f_n, s_n = h.compute()
    mode, fb, dev = h.sur_state()
    fdev = h.last_force_dev()
    h.close()
    print(f"  judge: mode {mode}, |dE|/atom {1e3 * dev:.4f} meV, force RMS dev {fdev:.3e} eV/A", flush=True)
    assert dev >= 0 and fdev >= 0
    if mode == d3_lib.D3X_ATM_SURROGATE:
        assert fb == 0 and dev <= 1e-3 and fdev <= 5e-3
        assert abs(e_n - e_x) / n == pytest.approx(dev, rel=1e-6, abs=1e-9)
        assert float(np.sqrt(np.mean((f_n - f_x) ** 2))) == pytest.approx(fdev, rel=1e-3)
    else:
        assert fb == 1 and (dev > 1e-3 or fdev > 5e-3)
        assert abs(e_n - e_x) <= 1e-6 * abs(e_x) and np.max(np.abs(f_n - f_x)) <= 1e-5
