"""Edge cases through the C ABI on the GPU: an empty input, a batch with empty and one-atom
cells among others, a single atom that interacts only with its own periodic images (every
triangle index-degenerate), a length above max_length (rejected, handle still usable), and a
centre with more forward three-body neighbours than the scratch capacity (grown and evaluated
again).  The FP64 engine is held to the oracle at the parity bar wherever there is something to
compare."""

import ctypes as C

import numpy as np
import pytest

import d3_lib
import d3_oracle
import d3_structures as S

pytestmark = pytest.mark.gpu

KW = dict(cutoff=60.0, cn_cutoff=30.0, atm_cutoff=25.0, damping=1, functional=0)


def _assert_parity(out, o):
    e, f, s = out
    assert abs(e - o["energy"]) <= 1e-12 * abs(o["energy"])
    assert np.max(np.abs(f - o["forces"])) <= 1e-9
    assert np.max(np.abs(s - o["stress"])) <= 1e-12 * np.max(np.abs(o["stress"])) + 1e-12


@pytest.mark.parametrize("precision", [d3_lib.D3X_FP64, d3_lib.D3X_FP32])
def test_empty_input(precision):
    print(f"case empty precision={precision}", flush=True)
    h = d3_lib.Handle([6, 8], 16, precision=precision, **KW)
    h.set_atoms(np.zeros((0, 3)), np.zeros(0, np.uint16))
    h.set_cell(np.eye(3) * 10.0)
    e, f, s = h.compute()
    h.close()
    assert e == 0.0 and f.shape == (0, 3) and np.all(s == 0.0)


def test_batch_with_empty_and_single_atom_cells(table):
    print("case batch with empty and one-atom cells", flush=True)
    rng = np.random.Generator(np.random.PCG64(21))
    cells, pos, Z, counts = [], [], [], []
    for n, edge in ((7, 8.0), (0, 9.0), (1, 4.5), (5, 7.5)):
        cell = np.eye(3) * edge + rng.uniform(-0.4, 0.4, (3, 3))
        p, z = S.random_packing(n, cell, rng, elements=np.array([1, 6, 8], np.uint16)) if n else (np.zeros((0, 3)),
                                                                                                    np.zeros(0, np.uint16))
        cells.append(cell)
        pos.append(p)
        Z.append(z)
        counts.append(n)
    h = d3_lib.Handle([1, 6, 8], sum(counts), **KW)
    e, f, s = h.compute_batch(counts, np.concatenate(pos), np.concatenate(Z), np.array(cells))
    h.close()
    a = 0
    for k, n in enumerate(counts):
        if n == 0:
            assert e[k] == 0.0 and np.all(s[k] == 0.0)
            continue
        o = d3_oracle.compute(pos[k], Z[k], cells[k], table, **KW)
        _assert_parity((e[k], f[a:a + n], s[k]), o)
        a += n


def test_single_atom_with_its_images(table):
    """One Cu atom in a 3.6 A triclinic cell: every pair and triangle is made of its own images."""
    print("case single atom", flush=True)
    cell = np.array([[3.6, 0.0, 0.0], [0.9, 3.5, 0.0], [-0.4, 0.7, 3.3]])
    pos = np.array([[0.3, 0.2, 0.1]])
    Z = np.array([29], np.uint16)
    o = d3_oracle.compute(pos, Z, cell, table, **KW)
    h = d3_lib.Handle(Z, 1, **KW)
    h.set_atoms(pos, Z)
    h.set_cell(cell)
    out = h.compute()
    h.close()
    _assert_parity(out, o)
    assert np.max(np.abs(out[1])) <= 1e-9  # one atom: no net force


def test_length_above_max_length_is_rejected(table):
    print("case max_length", flush=True)
    pos, Z, cell = S.si_diamond(nrep=1, jitter=0.02, seed=3)  # 8 atoms
    h = d3_lib.Handle(Z, 8, **KW)
    with pytest.raises(d3_lib.D3Error) as err:
        h.set_atoms(np.concatenate([pos, pos[:1] + 0.5]), np.concatenate([Z, Z[:1]]))
    assert err.value.code & d3_lib.D3X_ERR_INPUT
    h.set_atoms(pos, Z)  # the handle is still usable
    h.set_cell(cell)
    out = h.compute()
    h.close()
    _assert_parity(out, d3_oracle.compute(pos, Z, cell, table, **KW))


def hub_and_cap(radii=(12.0, 16.0, 20.0), zmin=8.0, per_a2=0.459, seed=2):
    """A hub atom in a 200 A cell and about 1200 atoms on the parts of three spheres around it (radii
    in A) that lie at least zmin A above it: zmin exceeds the bin width (15 bohr, 7.9 A), so every
    such atom sits in a bin layer above the hub's and is a forward three-body neighbour of it,
    while a cap atom has far fewer forward neighbours."""
    rng = np.random.Generator(np.random.PCG64(seed))
    pts = []
    for R in radii:
        n = int(per_a2 * 4.0 * np.pi * R * R)
        k = np.arange(n) + 0.5
        zc = 1.0 - 2.0 * k / n
        ph = np.pi * (1.0 + 5.0 ** 0.5) * k + rng.uniform(0.0, 2.0 * np.pi)  # turned about z
        rr = np.sqrt(1.0 - zc * zc)
        v = R * np.stack([rr * np.cos(ph), rr * np.sin(ph), zc], axis=1)
        pts.append(v[v[:, 2] >= zmin])
    hub = np.array([0.3, 0.3, 0.3])
    pos = np.vstack([hub, hub + np.vstack(pts)])
    Z = np.full(len(pos), 6, np.uint16)
    return pos, Z, np.eye(3) * 200.0


def test_three_body_capacity_overflow_is_retried(table):
    print("case three-body capacity overflow", flush=True)
    kw = dict(KW, atm_cutoff=40.0)
    pos, Z, cell = hub_and_cap()
    nfwd = len(Z) - 1
    # the first capacity estimate (1.5 x the mean forward count at this density, + 1024, rounded
    # up to 64) is below the hub's forward neighbours, so the first pass overflows
    dens = len(Z) / (200.0 / 0.52917721067) ** 3
    first = int(np.ceil((1.5 * 0.5 * dens * 4.18879020478639 * 40.0 ** 3 + 1024) / 64) * 64)
    assert first < nfwd
    o = d3_oracle.compute(pos, Z, cell, table, **kw)
    h = d3_lib.Handle(Z, len(Z), **kw)
    h.set_atoms(pos, Z)
    h.set_cell(cell)
    out = h.compute()
    cap = h.atm_capacity()
    h.close()
    assert cap >= nfwd
    _assert_parity(out, o)


def test_single_cell_calls_after_a_batch(table):
    """A handle that evaluated a batch: the reference's compute entry refuses it (one energy and
    one stress would not hold the batch's results), and set_cell alone does not turn the batch's
    atoms into one cell -- set_atoms is required again, after which the handle is a one-cell
    handle at the parity bar."""
    print("case single-cell calls after a batch", flush=True)
    rng = np.random.Generator(np.random.PCG64(31))
    cells, pos, Z, counts = [], [], [], []
    for n, edge in ((6, 8.0), (9, 9.0)):
        cell = np.eye(3) * edge + rng.uniform(-0.4, 0.4, (3, 3))
        p, z = S.random_packing(n, cell, rng, elements=np.array([1, 6, 8], np.uint16))
        cells.append(cell)
        pos.append(p)
        Z.append(z)
        counts.append(n)
    h = d3_lib.Handle([1, 6, 8], sum(counts), **KW)
    h.compute_batch(counts, np.concatenate(pos), np.concatenate(Z), np.array(cells))
    lib = d3_lib.load()
    E = C.c_float(0.0)
    F = np.zeros((sum(counts), 3), np.float32)
    Sg = np.zeros(9, np.float32)
    st = lib.compute_dispersion_energy_from_handle_status(h.h, C.byref(E), d3_lib.ptr(F, C.c_float),
                                                          d3_lib.ptr(Sg, C.c_float))
    assert st == 1 and "batch" in d3_lib.last_error()
    h.set_cell(cells[1])
    with pytest.raises(d3_lib.D3Error) as err:
        h.compute()
    assert err.value.code & d3_lib.D3X_ERR_STATE
    h.set_atoms(pos[1], Z[1])
    out = h.compute()
    h.close()
    _assert_parity(out, d3_oracle.compute(pos[1], Z[1], cells[1], table, **KW))
