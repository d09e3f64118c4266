"""Structure files for the D3 engine: LASP .arc and extended XYZ, read as (positions in
Angstrom, atomic numbers, cell rows = lattice vectors in Angstrom).

.arc follows the reference's reader (src/parser.cpp:73-176): a structure starts at a line
"Energy <n> <x> <energy> [symmetry]", its cell is "PBC a b c alpha beta gamma" (lengths in
Angstrom, angles in degrees, converted with a along x and b in the xy plane as in
src/modules.h:253-277), and its atoms are lines of exactly 10 fields "symbol x y z ...".
Extended XYZ: per frame a count line, a comment line with Lattice="..." (and optionally
Properties=..., which locates the species and pos columns), then one line per atom.
"""

from __future__ import annotations

import numpy as np

SYMBOLS = ("H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar K Ca Sc Ti V Cr Mn Fe Co Ni Cu Zn Ga Ge As Se "
           "Br Kr Rb Sr Y Zr Nb Mo Tc Ru Rh Pd Ag Cd In Sn Sb Te I Xe Cs Ba La Ce Pr Nd Pm Sm Eu Gd Tb Dy "
           "Ho Er Tm Yb Lu Hf Ta W Re Os Ir Pt Au Hg Tl Pb Bi Po At Rn Fr Ra Ac Th Pa U Np Pu Am Cm Bk Cf "
           "Es Fm Md No Lr Rf Db Sg Bh Hs Mt Ds Rg Cn Nh Fl Mc Lv Ts Og").split()
Z_OF = {s: i + 1 for i, s in enumerate(SYMBOLS)}


def cell_from_parameters(a, b, c, alpha, beta, gamma) -> np.ndarray:
    """Lattice vectors (rows) from lengths and angles in degrees: a along x, b in the xy plane."""
    ca, cb, cg = np.cos(np.radians([alpha, beta, gamma]))
    sg = np.sin(np.radians(gamma))
    cz = np.sqrt(max(0.0, 1.0 - ca * ca - cb * cb - cg * cg + 2.0 * ca * cb * cg)) / sg
    return np.array([[a, 0.0, 0.0], [b * cg, b * sg, 0.0], [c * cb, c * (ca - cb * cg) / sg, c * cz]])


def cell_parameters(cell) -> tuple:
    """(a, b, c, alpha, beta, gamma) of lattice vectors in rows."""
    A, B, Cv = np.asarray(cell, float)
    la, lb, lc = np.linalg.norm(A), np.linalg.norm(B), np.linalg.norm(Cv)
    ang = lambda u, v, nu, nv: np.degrees(np.arccos(np.clip(u @ v / (nu * nv), -1, 1)))  # noqa: E731
    return la, lb, lc, ang(B, Cv, lb, lc), ang(A, Cv, la, lc), ang(A, B, la, lb)


def read_extxyz(path: str) -> list:
    frames = []
    with open(path) as f:
        lines = f.read().splitlines()
    k = 0
    while k < len(lines):
        if not lines[k].strip():
            k += 1
            continue
        n = int(lines[k].split()[0])
        comment = lines[k + 1]
        cell = None
        if 'Lattice="' in comment:
            s = comment.split('Lattice="', 1)[1].split('"', 1)[0]
            cell = np.array([float(x) for x in s.split()]).reshape(3, 3)
        col_sp, col_pos = 0, 1
        if "Properties=" in comment:
            spec = comment.split("Properties=", 1)[1].split()[0].split(":")
            col = 0
            for i in range(0, len(spec) - 2, 3):
                w = int(spec[i + 2])
                if spec[i] == "species":
                    col_sp = col
                if spec[i] == "pos":
                    col_pos = col
                col += w
        Z, pos = [], []
        for line in lines[k + 2:k + 2 + n]:
            t = line.split()
            Z.append(Z_OF[t[col_sp]])
            pos.append([float(x) for x in t[col_pos:col_pos + 3]])
        frames.append((np.array(pos), np.array(Z, np.uint16), cell))
        k += 2 + n
    return frames


def read_arc(path: str) -> list:
    frames, cur = [], None
    with open(path) as f:
        for line in f:
            t = line.split()
            if len(t) == 10 and t[0] in Z_OF and cur is not None:
                cur[1].append(Z_OF[t[0]])
                cur[0].append([float(x) for x in t[1:4]])
            elif len(t) == 7 and t[0] == "PBC" and cur is not None:
                cur[2] = cell_from_parameters(*[float(x) for x in t[1:7]])
            elif len(t) in (4, 5) and t[0] == "Energy":
                if cur is not None:
                    frames.append(cur)
                cur = [[], [], None]
    if cur is not None:
        frames.append(cur)
    return [(np.array(p), np.array(z, np.uint16), c) for p, z, c in frames]


def write_arc(path: str, frames) -> None:
    """Writes (pos, Z, cell) frames as .arc; positions are re-expressed in the cell orientation
    that the PBC line implies (a along x, b in the xy plane)."""
    with open(path, "w") as f:
        f.write("!BIOSYM archive 2\nPBC=ON\n")
        for k, (pos, Z, cell) in enumerate(frames):
            par = cell_parameters(cell)
            std = cell_from_parameters(*par)
            frac = np.asarray(pos, float) @ np.linalg.inv(np.asarray(cell, float))
            p = frac @ std
            f.write(f"          Energy     {k + 1}   0.0000  0.000000  C1\n")
            f.write("!DATE\n")
            f.write("PBC  " + "  ".join(f"{x:.9f}" for x in par) + "\n")
            for i, (z, r) in enumerate(zip(Z, p)):
                s = SYMBOLS[int(z) - 1]
                f.write(f"{s:<3} {r[0]:15.9f} {r[1]:15.9f} {r[2]:15.9f} CORE {i + 1:4d} {s} {s} 0.0000 {i + 1:4d}\n")
            f.write("end\nend\n")
