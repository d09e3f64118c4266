"""ctypes bindings of libd3.so (the C ABI of include/d3.h and include/d3x.h).

Used by d3_cffi.py when the cffi module is absent, and by Python callers that want the d3x
extensions (FP64 outputs, precision mode, batches, device-resident timing).  The library is
loaded from ``<this dir>/../build/libd3.so`` -- the same relative location the reference's
d3_cffi.py uses (src/d3_cffi.py:47-49) -- or from the path in the D3_LIBRARY environment
variable.  Loading fails loudly if the library has not been built.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("D3_LIBRARY", os.path.join(_HERE, "..", "build", "libd3.so"))

D3X_FP64 = 0
D3X_FP32 = 1
D3X_ERR_GENERIC = 1
D3X_ERR_OVERFLOW = 2
D3X_ERR_INPUT = 4
D3X_ERR_DEVICE = 8
D3X_ERR_STATE = 16

ZERO_DAMPING = 0
BJ_DAMPING = 1
FUNCTIONALS = {"PBE": 0, "PBE0": 1, "B3LYP": 2, "BLYP": 3, "BP86": 4, "REVPBE": 5, "CUSTOM": 99}

# Every symbol the headers declare: name -> (restype, argtypes)
_f = C.c_float
_fp = C.POINTER(C.c_float)
_d = C.c_double
_dp = C.POINTER(C.c_double)
_u16p = C.POINTER(C.c_uint16)
_u64 = C.c_uint64
_u64p = C.POINTER(C.c_uint64)
_vp = C.c_void_p
_ip = C.POINTER(C.c_int)

SIGNATURES = {
    # d3.h
    "compute_dispersion_energy": (None, [_fp, _u16p, _u64, _fp, _f, _f, C.c_int, C.c_int, _fp, _fp, _fp]),
    "init_d3_handle": (_vp, [_u16p, _u64, _u64, _f, _f, C.c_int, C.c_int]),
    "init_d3_handle_with_atm": (_vp, [_u16p, _u64, _u64, _f, _f, _f, C.c_int, C.c_int]),
    "set_atoms": (None, [_vp, _fp, _u16p, _u64]),
    "set_cell": (None, [_vp, _fp]),
    "free_d3_handle": (None, [_vp]),
    "clear_d3_handle": (None, [_vp]),
    "compute_dispersion_energy_from_handle": (None, [_vp, _fp, _fp, _fp]),
    "compute_dispersion_energy_from_handle_status": (C.c_uint16, [_vp, _fp, _fp, _fp]),
    # d3x.h
    "d3x_init": (_vp, [_u16p, _u64, _u64, _d, _d, _d, C.c_int, C.c_int, C.c_int]),
    "d3x_set_functional_params": (C.c_int, [_vp, _dp]),
    "d3x_set_atoms_f64": (C.c_int, [_vp, _dp, _u16p, _u64]),
    "d3x_set_cell_f64": (C.c_int, [_vp, _dp]),
    "d3x_compute_f64": (C.c_int, [_vp, _dp, _dp, _dp]),
    "d3x_get_atom_data": (C.c_int, [_vp, _dp, _dp, _dp]),
    "d3x_compute_device": (C.c_int, [_vp]),
    "d3x_fetch_results": (C.c_int, [_vp, _dp, _dp, _dp]),
    "d3x_get_counts": (C.c_int, [_vp, _dp]),
    "d3x_set_timing": (C.c_int, [_vp, C.c_int]),
    "d3x_phase_times": (C.c_int, [_vp, _dp]),
    "d3x_set_graph": (C.c_int, [_vp, C.c_int]),
    "d3x_set_tuning": (C.c_int, [_vp, C.c_double, C.c_int]),
    "d3x_get_tuning": (C.c_int, [_vp, _dp, _ip]),
    "d3x_get_atm_variant": (C.c_int, [_vp]),
    "d3x_get_c6_pairwise": (C.c_int, [_vp]),
    "d3x_get_atm_capacity": (C.c_int, [_vp]),
    "d3x_set_cn_smooth": (C.c_int, [_vp, C.c_int]),
    "d3x_set_skin": (C.c_int, [_vp, C.c_double]),
    "d3x_set_neighbour_lists": (C.c_int, [_vp, C.c_int]),
    "d3x_neighbour_stats": (C.c_int, [_vp, _ip, _ip]),
    "d3x_autotune": (C.c_double, [_vp, C.c_char_p]),
    "d3x_get_stream": (_vp, [_vp]),
    "d3x_set_stream": (C.c_int, [_vp, _vp]),
    "d3x_compute_batch_f64": (C.c_int, [_vp, _u64, _u64p, _dp, _u16p, _dp, _dp, _dp, _dp]),
    "d3x_last_status": (C.c_uint32, [_vp]),
    "d3x_last_error": (C.c_char_p, []),
    "d3x_table_load": (C.c_int, [C.c_char_p]),
    "d3x_table_name": (C.c_char_p, []),
    "d3x_table_shape": (C.c_int, [_ip, _ip, _ip, _ip]),
    "d3x_table_export": (C.c_int, [_fp, _fp, _fp, _fp]),
    # d3x_sur.h (device-buffer arguments are raw addresses: c_void_p)
    "d3x_sur_load": (C.c_int, [_vp, _fp]),
    "d3x_sur_load_file": (C.c_int, [_vp, C.c_char_p]),
    "d3x_set_atm_mode": (C.c_int, [_vp, C.c_int, C.c_int, C.c_double]),
    "d3x_sur_state": (C.c_int, [_vp, _ip, _ip, _dp]),
    "d3x_sur_set_force_tol": (C.c_int, [_vp, C.c_double]),
    "d3x_sur_last_force_dev": (C.c_double, [_vp]),
    "d3x_sur_features": (C.c_int, [_vp, _fp]),
    "d3x_sur_tangent": (C.c_int, [_vp, _dp, _fp]),
    "d3x_sur_packed_bytes": (C.c_size_t, []),
    "d3x_sur_train_bytes": (C.c_size_t, [_u64]),
    "d3x_sur_pack": (C.c_int, [_vp, _vp, _vp]),
    "d3x_sur_infer": (C.c_int, [_vp, _vp, _u64, _vp, _vp, _vp, _vp, _vp]),
    "d3x_sur_train_grad": (C.c_int, [_vp, _vp, C.c_int, _u64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "d3x_sur_train_grad_noalloc": (C.c_int, [_vp, _vp, C.c_int, _u64, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "d3x_sur_train_force_grad": (C.c_int, [_vp, _vp, C.c_int, _u64, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "d3x_sur_tie_grad": (C.c_int, [_vp, C.c_int, _vp]),
    "d3x_sur_adam": (C.c_int, [_vp, _vp, _vp, _vp, _f, _f, _f, _f, C.c_int, _vp]),
}

D3X_ATM_EXACT = 0
D3X_ATM_SURROGATE = 1
D3X_ATM_EXACT_FEATURES = 2
SUR_NPARAMS = 75020
SUR_NFEAT = 64

_lib = None


def load(path: str | None = None) -> C.CDLL:
    """Loads libd3.so once and declares every exported function's signature."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = os.path.abspath(path or LIB_PATH)
    if not os.path.exists(p):
        raise RuntimeError(f"libd3.so not found at {p}: build it with `make -C lasp-d3_amd` "
                           "(or __graft_entry__.build())")
    lib = C.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def last_error() -> str:
    s = load().d3x_last_error()
    return s.decode() if s else ""


def table() -> dict:
    """The active parameter table as numpy arrays (layout of csrc/d3_tables.hpp)."""
    lib = load()
    nel, nref, zoff, zmax = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    lib.d3x_table_shape(C.byref(nel), C.byref(nref), C.byref(zoff), C.byref(zmax))
    ne, nr = nel.value, nref.value
    c6ab = np.empty((ne, ne, nr, nr, 3), np.float32)
    r0ab = np.empty((ne, ne), np.float32)
    rcov = np.empty(ne, np.float32)
    r2r4 = np.empty(ne, np.float32)
    lib.d3x_table_export(ptr(c6ab, C.c_float), ptr(r0ab, C.c_float), ptr(rcov, C.c_float), ptr(r2r4, C.c_float))
    return dict(c6ab=c6ab, r0ab=r0ab, rcov=rcov, r2r4=r2r4, nel=ne, nref=nr, zoff=zoff.value, zmax=zmax.value,
                name=lib.d3x_table_name().decode())


def load_table(path: str | None) -> None:
    rc = load().d3x_table_load(path.encode() if path else None)
    if rc:
        raise RuntimeError(f"d3x_table_load failed: {last_error()}")


class D3Error(RuntimeError):
    def __init__(self, where: str, code: int):
        super().__init__(f"{where} failed with status {code}: {last_error()}")
        self.code = code


class Handle:
    """Owning wrapper of a d3x handle (FP64 in/out)."""

    def __init__(self, elements, max_length: int, cutoff: float = 94.8683298, cn_cutoff: float = 40.0,
                 atm_cutoff: float = 0.0, damping: int = BJ_DAMPING, functional: int = 0,
                 precision: int = D3X_FP64):
        self.lib = load()
        el = np.ascontiguousarray(np.unique(np.asarray(elements, dtype=np.uint16)))
        self.h = self.lib.d3x_init(ptr(el, C.c_uint16), len(el), int(max_length), float(cutoff), float(cn_cutoff),
                                   float(atm_cutoff), int(damping), int(functional), int(precision))
        if not self.h:
            raise RuntimeError(f"d3x_init failed: {last_error()}")
        self.n = 0

    def _check(self, rc: int, where: str):
        if rc:
            raise D3Error(where, rc)

    def set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float64)
        self._check(self.lib.d3x_set_functional_params(self.h, ptr(p, C.c_double)), "d3x_set_functional_params")

    def set_atoms(self, coords, elements):
        c = np.ascontiguousarray(coords, dtype=np.float64).reshape(-1, 3)
        z = np.ascontiguousarray(elements, dtype=np.uint16)
        self._check(self.lib.d3x_set_atoms_f64(self.h, ptr(c, C.c_double), ptr(z, C.c_uint16), len(z)),
                    "d3x_set_atoms_f64")
        self.n = len(z)

    def set_cell(self, cell):
        c = np.ascontiguousarray(cell, dtype=np.float64).reshape(9)
        self._check(self.lib.d3x_set_cell_f64(self.h, ptr(c, C.c_double)), "d3x_set_cell_f64")

    def compute(self):
        e = np.zeros(1)
        f = np.zeros((self.n, 3))
        s = np.zeros(9)
        self._check(self.lib.d3x_compute_f64(self.h, ptr(e, C.c_double), ptr(f, C.c_double), ptr(s, C.c_double)),
                    "d3x_compute_f64")
        return float(e[0]), f, s.reshape(3, 3)

    def atom_data(self):
        cn, g, ea = np.zeros(self.n), np.zeros(self.n), np.zeros(self.n)
        self._check(self.lib.d3x_get_atom_data(self.h, ptr(cn, C.c_double), ptr(g, C.c_double),
                                               ptr(ea, C.c_double)), "d3x_get_atom_data")
        return cn, g, ea

    def compute_device(self):
        self._check(self.lib.d3x_compute_device(self.h), "d3x_compute_device")

    def fetch(self):
        e = np.zeros(1)
        f = np.zeros((self.n, 3))
        s = np.zeros(9)
        self._check(self.lib.d3x_fetch_results(self.h, ptr(e, C.c_double), ptr(f, C.c_double), ptr(s, C.c_double)),
                    "d3x_fetch_results")
        return float(e[0]), f, s.reshape(3, 3)

    def compute_batch(self, counts, coords, elements, cells):
        counts = np.ascontiguousarray(counts, dtype=np.uint64)
        c = np.ascontiguousarray(coords, dtype=np.float64).reshape(-1, 3)
        z = np.ascontiguousarray(elements, dtype=np.uint16)
        cl = np.ascontiguousarray(cells, dtype=np.float64).reshape(-1, 9)
        ns = len(counts)
        e = np.zeros(ns)
        f = np.zeros((len(z), 3))
        s = np.zeros((ns, 9))
        self._check(self.lib.d3x_compute_batch_f64(self.h, ns, ptr(counts, C.c_uint64), ptr(c, C.c_double),
                                                   ptr(z, C.c_uint16), ptr(cl, C.c_double), ptr(e, C.c_double),
                                                   ptr(f, C.c_double), ptr(s, C.c_double)), "d3x_compute_batch_f64")
        self.n = len(z)
        return e, f, s.reshape(ns, 3, 3)

    def counts(self, nsys: int = 1) -> np.ndarray:
        c = np.zeros((nsys, 3))
        self._check(self.lib.d3x_get_counts(self.h, ptr(c, C.c_double)), "d3x_get_counts")
        return c

    def set_tuning(self, bin_width: float = 0.0, atm_blocks: int = 0):
        self._check(self.lib.d3x_set_tuning(self.h, float(bin_width), int(atm_blocks)), "d3x_set_tuning")

    def tuning(self):
        w, b = C.c_double(), C.c_int()
        self._check(self.lib.d3x_get_tuning(self.h, C.byref(w), C.byref(b)), "d3x_get_tuning")
        return w.value, b.value

    def atm_variant(self) -> int:
        """Three-body kernel variant in use (bit 0: k-side sums in LDS, bit 1: damping rows)."""
        v = self.lib.d3x_get_atm_variant(self.h)
        if v < 0:
            raise D3Error("d3x_get_atm_variant", v)
        return v

    def c6_pairwise(self) -> bool:
        """True when C6 is interpolated per pair (a table whose CN references do not factorise
        per element, or D3_C6_PAIRWISE=1 at init)."""
        return self.lib.d3x_get_c6_pairwise(self.h) == 1

    def atm_capacity(self) -> int:
        """Forward-neighbour capacity per centre of the three-body scratch."""
        return int(self.lib.d3x_get_atm_capacity(self.h))

    def set_cn_smooth(self, on: bool = True):
        """The smooth CN cutoff (tanh^3(rc_CN - r) on every CN term)."""
        self._check(self.lib.d3x_set_cn_smooth(self.h, int(bool(on))), "d3x_set_cn_smooth")

    def set_skin(self, skin_bohr: float):
        """Keep the neighbour state across evaluations while atoms move less than skin / 2."""
        self._check(self.lib.d3x_set_skin(self.h, float(skin_bohr)), "d3x_set_skin")

    def set_neighbour_lists(self, mode: int):
        """Three-body forward lists: -1 automatic, 0 never, 1 always."""
        self._check(self.lib.d3x_set_neighbour_lists(self.h, int(mode)), "d3x_set_neighbour_lists")

    def neighbour_stats(self):
        """(evaluations that rebuilt the neighbour state, evaluations that reused it)"""
        a, b = C.c_int(), C.c_int()
        self._check(self.lib.d3x_neighbour_stats(self.h, C.byref(a), C.byref(b)), "d3x_neighbour_stats")
        return a.value, b.value

    def autotune(self, cache_path: str | None = None) -> float:
        t = self.lib.d3x_autotune(self.h, cache_path.encode() if cache_path else None)
        if t < 0:
            raise D3Error("d3x_autotune", -1)
        return t

    def set_graph(self, on: bool):
        self._check(self.lib.d3x_set_graph(self.h, int(on)), "d3x_set_graph")

    def set_timing(self, on: bool):
        self._check(self.lib.d3x_set_timing(self.h, int(on)), "d3x_set_timing")

    PHASES = ("bin_sort", "cn", "two_body", "atm", "chain", "reduce")

    def phase_times(self):
        ms = np.zeros(6)
        n = self.lib.d3x_phase_times(self.h, ptr(ms, C.c_double))
        if n < 0:
            raise D3Error("d3x_phase_times", -1)
        return dict(zip(self.PHASES, ms)), n

    def stream(self) -> int:
        return self.lib.d3x_get_stream(self.h) or 0

    # ---- three-body stand-in (include/d3x_sur.h)
    def sur_load(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.size != SUR_NPARAMS:
            raise ValueError(f"expected {SUR_NPARAMS} parameters, got {p.size}")
        self._check(self.lib.d3x_sur_load(self.h, ptr(p, C.c_float)), "d3x_sur_load")

    def sur_load_file(self, path: str):
        self._check(self.lib.d3x_sur_load_file(self.h, path.encode()), "d3x_sur_load_file")

    def set_atm_mode(self, mode: int, judge_every: int = 0, tol_ev_per_atom: float = 1e-3):
        self._check(self.lib.d3x_set_atm_mode(self.h, int(mode), int(judge_every), float(tol_ev_per_atom)),
                    "d3x_set_atm_mode")

    def sur_state(self):
        """(mode in use, fallback count, last judged |dE| / atoms in eV)"""
        m, f, d = C.c_int(), C.c_int(), C.c_double()
        self._check(self.lib.d3x_sur_state(self.h, C.byref(m), C.byref(f), C.byref(d)), "d3x_sur_state")
        return m.value, f.value, d.value

    def set_judge_force_tol(self, tol_rms_ev_per_A: float):
        """Force criterion of the judge (RMS of F_net - F_exact per cell, eV/A; <= 0: energies only)."""
        self._check(self.lib.d3x_sur_set_force_tol(self.h, float(tol_rms_ev_per_A)), "d3x_sur_set_force_tol")

    def last_force_dev(self) -> float:
        """RMS force deviation of the last judged evaluation (eV/A; negative before any)."""
        return float(self.lib.d3x_sur_last_force_dev(self.h))

    def features(self) -> np.ndarray:
        x = np.zeros((self.n, SUR_NFEAT), np.float32)
        self._check(self.lib.d3x_sur_features(self.h, ptr(x, C.c_float)), "d3x_sur_features")
        return x

    def sur_tangent(self, direction) -> np.ndarray:
        """U = J dir (atoms x 64) at the last evaluation's geometry; dir atoms x 3 per A."""
        d = np.ascontiguousarray(direction, dtype=np.float64).reshape(self.n, 3)
        u = np.zeros((self.n, SUR_NFEAT), np.float32)
        self._check(self.lib.d3x_sur_tangent(self.h, ptr(d, C.c_double), ptr(u, C.c_float)), "d3x_sur_tangent")
        return u

    def status(self) -> int:
        return int(self.lib.d3x_last_status(self.h))

    def close(self):
        if getattr(self, "h", None):
            self.lib.free_d3_handle(self.h)
            self.h = None

    def __del__(self):
        self.close()
