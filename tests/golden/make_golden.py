"""Writes the golden fixtures of this directory: FP64 oracle results (oracle/d3_oracle) for the
reference's example structure (tio2_192.extxyz, a copy of the reference's
examples/structures/192.extxyz) with the reference driver's settings (src/main.cpp:60-74: cutoff
60 bohr, CN and ATM cutoffs 40 bohr, zero damping, PBE) and for the seeded parity systems of
tests/systems.py.  Inputs and expected outputs go to golden.npz (float64).

usage: python tests/golden/make_golden.py
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for sub in ("lasp-d3_amd/src", "lasp-d3_amd/tables", "oracle", "tests"):
    sys.path.insert(0, os.path.join(ROOT, sub))

import d3_io  # noqa: E402
import d3_lib  # noqa: E402
import d3_oracle  # noqa: E402
import systems  # noqa: E402


def cases():
    (pos, Z, cell), = d3_io.read_extxyz(os.path.join(HERE, "tio2_192.extxyz"))
    drv = dict(cutoff=60.0, cn_cutoff=40.0, damping=0, functional=0)
    yield "tio2_192", pos, Z, cell, dict(drv, atm_cutoff=0.0)
    yield "tio2_192_atm", pos, Z, cell, dict(drv, atm_cutoff=40.0)
    for name, s in (("si64_bj_atm", systems.si64(seed=1)), ("triclinic24", systems.triclinic_mixed(n=24, seed=3)),
                    ("water16", systems.water_small(16, seed=5)), ("skewed_zero", systems.skewed_zero(seed=7))):
        yield name, s["pos"], s["Z"], s["cell"], dict(cutoff=s["cutoff"], cn_cutoff=s["cn_cutoff"],
                                                        atm_cutoff=s["atm"], damping=s["damping"],
                                                        functional=s.get("functional", 0))


def main():
    table = d3_lib.table()
    out = {}
    for name, pos, Z, cell, kw in cases():
        o = d3_oracle.compute(pos, Z, cell, table, **kw)
        out[f"{name}/pos"] = np.asarray(pos, np.float64)
        out[f"{name}/Z"] = np.asarray(Z, np.int64)
        out[f"{name}/cell"] = np.asarray(cell, np.float64)
        out[f"{name}/settings"] = np.array([kw["cutoff"], kw["cn_cutoff"], kw["atm_cutoff"], kw["damping"],
                                            kw["functional"]], np.float64)
        for k in ("energy", "forces", "stress", "cn", "dedcn", "atom_energy"):
            out[f"{name}/{k}"] = np.asarray(o[k], np.float64)
        print(name, len(Z), o["energy"], flush=True)
    np.savez_compressed(os.path.join(HERE, "golden.npz"), **out)


if __name__ == "__main__":
    main()
