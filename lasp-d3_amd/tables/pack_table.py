#!/usr/bin/env python3
"""Packs real D3 parameter tables into the engine's table-file format (csrc/d3_tables.hpp).

Sources:
  * the reference's generated headers src/gen_constants_original.h (95 x 5 references, slot = Z)
    or src/gen_constants_extended.h (103 x 7, slot = Z - 1), selected in the reference by
    USE_EXTENDED_PARAMETERS (src/constants_include.h:9-32): C initialisers of c6ab_ref, r0ab,
    rcov and r2r4, read as flat lists of numbers in declaration order;
  * an .npz with arrays c6ab (NEL, NEL, NREF, NREF, 3), r0ab (NEL, NEL), rcov (NEL), r2r4 (NEL).

    python pack_table.py gen_constants_original.h d3_original.d3tab
    python pack_table.py tables.npz d3_custom.d3tab --extended

The file is then selected with d3x_table_load("d3_original.d3tab") before creating handles.
rcov must already include the k2 = 4/3 factor and r2r4 must be the per-element factor with
C8 = 3 C6 r2r4_A r2r4_B, as in the reference's tables.
"""

from __future__ import annotations

import argparse
import re
import struct

import numpy as np

MAGIC = b"D3TAB001"


def write_table(path, c6ab, r0ab, rcov, r2r4, zoff: int, zmax: int) -> None:
    c6ab = np.ascontiguousarray(c6ab, dtype=np.float32)
    nel, nel2, nref, nref2, three = c6ab.shape
    if nel != nel2 or nref != nref2 or three != 3:
        raise ValueError(f"c6ab has shape {c6ab.shape}; expected (NEL, NEL, NREF, NREF, 3)")
    r0ab = np.ascontiguousarray(r0ab, dtype=np.float32).reshape(nel, nel)
    rcov = np.ascontiguousarray(rcov, dtype=np.float32).reshape(nel)
    r2r4 = np.ascontiguousarray(r2r4, dtype=np.float32).reshape(nel)
    with open(path, "wb") as f:
        f.write(MAGIC)
        f.write(struct.pack("<4i", nel, nref, zoff, zmax))
        for a in (c6ab, r0ab, rcov, r2r4):
            f.write(a.tobytes())


_NUM = re.compile(r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?[fF]?")


def _initializer(src: str, name: str) -> np.ndarray:
    m = re.search(r"\b" + re.escape(name) + r"\s*(?:\[[^\]]*\]\s*)+=\s*\{", src)
    if not m:
        raise ValueError(f"array {name} not found")
    depth, i = 1, m.end()
    while depth and i < len(src):
        depth += {"{": 1, "}": -1}.get(src[i], 0)
        i += 1
    body = re.sub(r"//[^\n]*|/\*.*?\*/", "", src[m.end():i - 1], flags=re.S)
    return np.array([float(t.rstrip("fF")) for t in _NUM.findall(body)], dtype=np.float64)


def read_header(path: str, extended: bool):
    nel, nref = (103, 7) if extended else (95, 5)
    src = open(path).read()
    c6 = _initializer(src, "c6ab_ref")
    if c6.size != nel * nel * nref * nref * 3:
        raise ValueError(f"c6ab_ref has {c6.size} values, expected {nel * nel * nref * nref * 3}")
    r0 = _initializer(src, "r0ab").reshape(nel, nel)
    rc = _initializer(src, "rcov")
    rr = _initializer(src, "r2r4")
    return c6.reshape(nel, nel, nref, nref, 3), r0, rc, rr


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("src")
    ap.add_argument("dst")
    ap.add_argument("--extended", action="store_true", help="103 x 7 layout with slot = Z - 1")
    a = ap.parse_args()
    if a.src.endswith(".npz"):
        d = np.load(a.src, allow_pickle=False)
        c6, r0, rc, rr = d["c6ab"], d["r0ab"], d["rcov"], d["r2r4"]
        extended = a.extended or c6.shape[0] == 103
    else:
        extended = a.extended or "extended" in a.src
        c6, r0, rc, rr = read_header(a.src, extended)
    zoff, zmax = (1, 103) if extended else (0, 94)
    write_table(a.dst, c6, r0, rc, rr, zoff, zmax)
    print(f"wrote {a.dst}: NEL={c6.shape[0]} NREF={c6.shape[2]} zoff={zoff} zmax={zmax}")


if __name__ == "__main__":
    main()
