"""Seeded structure generators for the benchmark and test workloads (Angstrom, rows = lattice
vectors).  numpy's PCG64 generator is platform-independent, so the same seed gives bit-identical
inputs on every machine.

  si_diamond      diamond-cubic Si supercell, Gaussian jitter           (512 atoms at 4x4x4)
  water_box       rigid H2O at a given density, random orientations       (4096 molecules)
  triclinic_batch many triclinic cells of uneven size, 6 elements         (~300k atoms at 256)
  fcc_cu          fcc Cu supercell, Gaussian jitter                       (1,000,188 atoms at 63^3)
  random_cell     random 6-element packing at 0.08 atoms/A^3             (100k atoms)
"""

from __future__ import annotations

import math

import numpy as np

SIX_ELEMENTS = np.array([1, 6, 7, 8, 14, 29], dtype=np.uint16)


def si_diamond(nrep: int = 4, a: float = 5.431, jitter: float = 0.02, seed: int = 0):
    basis = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0],
                      [.25, .25, .25], [.25, .75, .75], [.75, .25, .75], [.75, .75, .25]])
    g = np.stack(np.meshgrid(np.arange(nrep), np.arange(nrep), np.arange(nrep), indexing="ij"), -1).reshape(-1, 3)
    frac = (g[:, None, :] + basis[None]).reshape(-1, 3)
    pos = frac * a
    rng = np.random.Generator(np.random.PCG64(seed))
    pos = pos + rng.normal(0.0, jitter, pos.shape)
    cell = np.eye(3) * a * nrep
    return pos, np.full(len(pos), 14, np.uint16), cell


def fcc_cu(nrep: int = 63, a: float = 3.615, jitter: float = 0.05, seed: int = 0):
    basis = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]])
    r = np.arange(nrep)
    g = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3)
    frac = (g[:, None, :] + basis[None]).reshape(-1, 3)
    pos = frac * a
    rng = np.random.Generator(np.random.PCG64(seed))
    pos = pos + rng.normal(0.0, jitter, pos.shape)
    return pos, np.full(len(pos), 29, np.uint16), np.eye(3) * a * nrep


def _random_rotations(rng, n):
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1),
    ], 1)


def water_box(n_mol: int = 4096, density: float = 1.0, seed: int = 0, d_oo: float = 2.5):
    """Rigid H2O (O-H 0.9572 A, H-O-H 104.52 deg) at `density` g/cm^3 in a cubic cell; oxygen
    positions uniform with rejection below d_oo (periodic minimum image), uniform rotations."""
    mass = 18.015 / 6.02214076e23  # g per molecule
    edge = (n_mol * mass / density * 1e24) ** (1.0 / 3.0)
    rng = np.random.Generator(np.random.PCG64(seed))
    ng = max(1, int(edge // d_oo))
    gw = edge / ng
    grid: dict = {}
    O = np.empty((n_mol, 3))
    k = 0
    while k < n_mol:
        p = rng.uniform(0.0, edge, 3)
        c = (p // gw).astype(int) % ng
        ok = True
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for q in grid.get(((c[0] + dx) % ng, (c[1] + dy) % ng, (c[2] + dz) % ng), ()):
                        d = p - O[q]
                        d -= edge * np.round(d / edge)
                        if d @ d < d_oo * d_oo:
                            ok = False
                            break
                    if not ok:
                        break
                if not ok:
                    break
        if ok:
            O[k] = p
            grid.setdefault(tuple(c), []).append(k)
            k += 1
    th = math.radians(104.52) / 2
    local = np.array([[0, 0, 0], [math.sin(th), math.cos(th), 0], [-math.sin(th), math.cos(th), 0]]) * 0.9572
    R = _random_rotations(rng, n_mol)
    pos = O[:, None, :] + np.einsum("nij,aj->nai", R, local)
    Z = np.tile(np.array([8, 1, 1], np.uint16), n_mol)
    return pos.reshape(-1, 3), Z, np.eye(3) * edge


def triclinic_cell(volume: float, angles_deg) -> np.ndarray:
    al, be, ga = (math.radians(x) for x in angles_deg)
    ca, cb, cg, sg = math.cos(al), math.cos(be), math.cos(ga), math.sin(ga)
    v1 = math.sqrt(max(1e-12, 1 - ca * ca - cb * cb - cg * cg + 2 * ca * cb * cg))
    L = (volume / v1) ** (1.0 / 3.0)
    a1 = [L, 0, 0]
    a2 = [L * cg, L * sg, 0]
    a3x = L * cb
    a3y = L * (ca - cb * cg) / sg
    a3z = math.sqrt(max(0.0, L * L - a3x * a3x - a3y * a3y))
    return np.array([a1, a2, [a3x, a3y, a3z]])


def random_packing(n: int, cell: np.ndarray, rng, elements=SIX_ELEMENTS, d_min: float = 1.0):
    """Random sequential packing of n atoms in a periodic (triclinic) cell with minimum-image
    distance >= d_min; elements uniform from `elements`."""
    inv = np.linalg.inv(cell)
    heights = 1.0 / np.linalg.norm(inv, axis=0)
    ng = np.maximum(1, (heights // max(d_min, 1e-9)).astype(int))
    frac = np.empty((n, 3))
    grid: dict = {}
    k = 0
    tries = 0
    while k < n:
        tries += 1
        if tries > 200 * n + 1000:
            raise RuntimeError("random packing failed (density too high)")
        s = rng.uniform(0.0, 1.0, 3)
        c = (s * ng).astype(int) % ng
        ok = True
        cand = []
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    cand.extend(grid.get(((c[0] + dx) % ng[0], (c[1] + dy) % ng[1], (c[2] + dz) % ng[2]), ()))
        if cand:
            ds = frac[cand] - s
            ds -= np.round(ds)
            d = ds @ cell
            ok = bool(np.all(np.einsum("ij,ij->i", d, d) >= d_min * d_min))
        if ok:
            frac[k] = s
            grid.setdefault(tuple(c), []).append(k)
            k += 1
    Z = rng.choice(elements, size=n).astype(np.uint16)
    return frac @ cell, Z


def triclinic_batch(n_struct: int = 256, seed: int = 0, n_min: int = 50, n_max: int = 5000,
                    density: float = 0.08):
    """Cells with log-uniform atom counts in [n_min, n_max], angles U(70, 110) deg, volume
    N / density, elements uniform over {H, C, N, O, Si, Cu}, d_min = 1 A packing."""
    rng = np.random.Generator(np.random.PCG64(seed))
    counts = np.exp(rng.uniform(math.log(n_min), math.log(n_max), n_struct)).astype(np.int64)
    pos, Z, cells = [], [], []
    for n in counts:
        ang = rng.uniform(70.0, 110.0, 3)
        cell = triclinic_cell(n / density, ang)
        p, z = random_packing(int(n), cell, rng)
        pos.append(p)
        Z.append(z)
        cells.append(cell)
    return counts.astype(np.uint64), np.concatenate(pos), np.concatenate(Z), np.stack(cells)


def random_cell(n: int = 100000, density: float = 0.08, seed: int = 0):
    rng = np.random.Generator(np.random.PCG64(seed))
    edge = (n / density) ** (1.0 / 3.0)
    cell = np.eye(3) * edge
    p, z = random_packing(n, cell, rng)
    return p, z, cell
