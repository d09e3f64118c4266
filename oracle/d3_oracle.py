"""ctypes wrapper of the FP64 oracle (oracle/libd3_oracle.so) -- test infrastructure only.

Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, as the
checker / CPU baseline; the product never loads it.  Units as in d3.h: Angstrom in, cutoffs in
bohr; E eV, F eV/Angstrom, stress eV/Angstrom^3, per-atom energies eV, dE/dCN Hartree.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# D3_ORACLE_LIBRARY: another build of the same source (the sanitizer build of oracle/Makefile)
LIB = os.environ.get("D3_ORACLE_LIBRARY", os.path.join(_HERE, "libd3_oracle.so"))
_lib = None

# {s6, s8_zero, s8_bj, sr6, sr8, a1, a2} of the reference's FUNCTIONAL_PARAMS (src/d3_types.h:46-54),
# as float32 values like the reference stores them.
FUNCTIONAL_PARAMS = np.array([
    [1.0, 0.722, 0.7875, 1.217, 1.0, 0.4289, 4.4407],
    [1.0, 0.928, 1.2177, 1.287, 1.0, 0.4145, 4.8593],
    [1.0, 1.703, 1.9889, 1.261, 1.0, 0.3981, 4.4211],
    [1.0, 1.682, 2.6996, 1.094, 1.0, 0.4298, 4.2359],
    [1.0, 1.683, 3.2822, 1.139, 1.0, 0.3946, 4.8516],
    [1.0, 1.010, 2.3550, 0.923, 1.0, 0.5238, 3.5016],
    [1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0],
], dtype=np.float32).astype(np.float64)


def functional_params(functional: int) -> np.ndarray:
    return FUNCTIONAL_PARAMS[6 if functional == 99 else functional].copy()


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            subprocess.run(["make", "-C", _HERE], check=True, capture_output=True)
        lib = C.CDLL(LIB)
        dp, fp = C.POINTER(C.c_double), C.POINTER(C.c_float)
        i32p, i64p, u64p = C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_uint64)
        lib.d3o_compute.restype = C.c_int
        lib.d3o_compute.argtypes = [dp, i32p, C.c_int64, dp, fp, fp, fp, fp, C.c_int32, C.c_int32, C.c_int32,
                                    C.c_double, C.c_double, C.c_double, C.c_int32, dp, dp, dp, dp, dp, dp, dp, u64p]
        lib.d3o_time_centers.restype = C.c_double
        lib.d3o_time_centers.argtypes = [dp, i32p, C.c_int64, dp, fp, fp, fp, fp, C.c_int32, C.c_int32, C.c_int32,
                                         C.c_double, C.c_double, C.c_double, C.c_int32, dp, i64p, C.c_int64, dp,
                                         dp, C.c_int32, u64p]
        lib.d3o_centers.restype = C.c_int
        lib.d3o_centers.argtypes = [dp, i32p, C.c_int64, dp, fp, fp, fp, fp, C.c_int32, C.c_int32, C.c_int32,
                                    C.c_double, C.c_double, C.c_double, C.c_int32, dp, i64p, C.c_int64, dp, dp,
                                    dp, dp, dp, dp]
        lib.d3o_max_threads.restype = C.c_int
        lib.d3o_set_cn_smooth.restype = None
        lib.d3o_set_cn_smooth.argtypes = [C.c_int32]
        _lib = lib
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def compute(coords, Z, cell, table: dict, cutoff=94.8683298, cn_cutoff=40.0, atm_cutoff=0.0, damping=1,
            functional=0, params=None, smooth_cn=False) -> dict:
    """Full FP64 evaluation.  table: dict with c6ab, r0ab, rcov, r2r4 (float32), nel, nref, zoff.
    smooth_cn: the tanh^3 smooth CN cutoff (the reference's SMOOTH_CUTOFF_CN build option)."""
    lib = load()
    lib.d3o_set_cn_smooth(1 if smooth_cn else 0)
    x = np.ascontiguousarray(coords, dtype=np.float64).reshape(-1, 3)
    z = np.ascontiguousarray(Z, dtype=np.int32)
    cl = np.ascontiguousarray(cell, dtype=np.float64).reshape(9)
    n = len(z)
    prm = np.ascontiguousarray(functional_params(functional) if params is None else params, dtype=np.float64)
    tb = {k: np.ascontiguousarray(table[k], dtype=np.float32) for k in ("c6ab", "r0ab", "rcov", "r2r4")}
    e = np.zeros(1)
    f = np.zeros((n, 3))
    s = np.zeros(9)
    cn, g, ea = np.zeros(n), np.zeros(n), np.zeros(n)
    counts = np.zeros(3, dtype=np.uint64)
    rc = lib.d3o_compute(_p(x, C.c_double), _p(z, C.c_int32), n, _p(cl, C.c_double), _p(tb["c6ab"], C.c_float),
                         _p(tb["r0ab"], C.c_float), _p(tb["rcov"], C.c_float), _p(tb["r2r4"], C.c_float),
                         table["nel"], table["nref"], table["zoff"], float(cutoff), float(cn_cutoff),
                         float(atm_cutoff), int(damping), _p(prm, C.c_double), _p(e, C.c_double),
                         _p(f, C.c_double), _p(s, C.c_double), _p(cn, C.c_double), _p(g, C.c_double),
                         _p(ea, C.c_double), _p(counts, C.c_uint64))
    lib.d3o_set_cn_smooth(0)
    if rc:
        raise ValueError(f"oracle rejected the input (code {rc})")
    return dict(energy=float(e[0]), forces=f, stress=s.reshape(3, 3), cn=cn, dedcn=g, atom_energy=ea,
                n2b=int(counts[0]), ncn=int(counts[1]), ntri=int(counts[2]))


def time_centers(coords, Z, cell, table: dict, centers, cn, dedcn, cutoff=94.8683298, cn_cutoff=40.0,
                 atm_cutoff=0.0, damping=1, functional=0, nthreads=0):
    """Wall time (s) of the four passes for the given centre atoms, and the pair counts."""
    lib = load()
    x = np.ascontiguousarray(coords, dtype=np.float64).reshape(-1, 3)
    z = np.ascontiguousarray(Z, dtype=np.int32)
    cl = np.ascontiguousarray(cell, dtype=np.float64).reshape(9)
    prm = functional_params(functional)
    tb = {k: np.ascontiguousarray(table[k], dtype=np.float32) for k in ("c6ab", "r0ab", "rcov", "r2r4")}
    cen = np.ascontiguousarray(centers, dtype=np.int64)
    cnv = np.ascontiguousarray(cn, dtype=np.float64)
    gv = np.ascontiguousarray(dedcn, dtype=np.float64)
    counts = np.zeros(3, dtype=np.uint64)
    t = lib.d3o_time_centers(_p(x, C.c_double), _p(z, C.c_int32), len(z), _p(cl, C.c_double),
                             _p(tb["c6ab"], C.c_float), _p(tb["r0ab"], C.c_float), _p(tb["rcov"], C.c_float),
                             _p(tb["r2r4"], C.c_float), table["nel"], table["nref"], table["zoff"], float(cutoff),
                             float(cn_cutoff), float(atm_cutoff), int(damping), _p(prm, C.c_double),
                             _p(cen, C.c_int64), len(cen), _p(cnv, C.c_double), _p(gv, C.c_double), int(nthreads),
                             _p(counts, C.c_uint64))
    return t, counts


def centers(coords, Z, cell, table: dict, centers, cn, dedcn, cutoff=94.8683298, cn_cutoff=40.0,
            atm_cutoff=0.0, damping=1, functional=0) -> dict:
    """Per-centre results for the given centre atoms of a large cell: cn from the CN pass; with the
    CN and total dE/dCN of every atom taken from cn / dedcn, the per-atom energy (eV), the
    two-body + three-body dE/dCN (Hartree) and the force (eV/Angstrom) of each centre."""
    lib = load()
    x = np.ascontiguousarray(coords, dtype=np.float64).reshape(-1, 3)
    z = np.ascontiguousarray(Z, dtype=np.int32)
    cl = np.ascontiguousarray(cell, dtype=np.float64).reshape(9)
    prm = functional_params(functional)
    tb = {k: np.ascontiguousarray(table[k], dtype=np.float32) for k in ("c6ab", "r0ab", "rcov", "r2r4")}
    cen = np.ascontiguousarray(centers, dtype=np.int64)
    cnv = np.ascontiguousarray(cn, dtype=np.float64)
    gv = np.ascontiguousarray(dedcn, dtype=np.float64)
    nc = len(cen)
    ocn, oe, og, of = np.zeros(nc), np.zeros(nc), np.zeros(nc), np.zeros((nc, 3))
    rc = lib.d3o_centers(_p(x, C.c_double), _p(z, C.c_int32), len(z), _p(cl, C.c_double), _p(tb["c6ab"], C.c_float),
                         _p(tb["r0ab"], C.c_float), _p(tb["rcov"], C.c_float), _p(tb["r2r4"], C.c_float),
                         table["nel"], table["nref"], table["zoff"], float(cutoff), float(cn_cutoff),
                         float(atm_cutoff), int(damping), _p(prm, C.c_double), _p(cen, C.c_int64), nc,
                         _p(cnv, C.c_double), _p(gv, C.c_double), _p(ocn, C.c_double), _p(oe, C.c_double),
                         _p(og, C.c_double), _p(of, C.c_double))
    if rc:
        raise ValueError(f"oracle rejected the input (code {rc})")
    return dict(cn=ocn, atom_energy=oe, dedcn=og, forces=of)


def max_threads() -> int:
    return int(load().d3o_max_threads())
