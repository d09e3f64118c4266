"""C ABI checks that need no GPU: every function declared in include/*.h is exported by
libd3.so and declared by the ctypes bindings; argument validation that happens before any device
call; the table functions (synthetic set, file round trip)."""

import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import d3_lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in ("d3.h", "d3x.h", "d3x_sur.h"):
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", src):
            name = m.group(1)
            if name.startswith(("d3x_", "init_d3", "set_", "free_d3", "clear_d3", "compute_")):
                names.add(name)
    return names


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) == 57  # 9 reference entry points + 31 extensions + 17 of the three-body stand-in
    out = subprocess.run(["nm", "-D", "--defined-only", d3_lib.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = names - exported
    assert not missing, missing
    lib = d3_lib.load()
    for n in names:
        getattr(lib, n)


def test_bindings_declare_every_symbol():
    assert declared_functions() == set(d3_lib.SIGNATURES)


def test_reference_entry_points_present():
    # the nine functions of the reference's src/d3.h:36-128
    ref = {"compute_dispersion_energy", "init_d3_handle", "init_d3_handle_with_atm", "set_atoms", "set_cell",
           "free_d3_handle", "clear_d3_handle", "compute_dispersion_energy_from_handle",
           "compute_dispersion_energy_from_handle_status"}
    assert ref <= declared_functions()


def test_invalid_init_is_rejected_before_device_use():
    lib = d3_lib.load()
    el = np.array([6, 200], np.uint16)  # 200 is beyond every table
    h = lib.init_d3_handle(d3_lib.ptr(el, C.c_uint16), 2, 10, 40.0, 40.0, 0, 0)
    assert not h
    assert "element" in d3_lib.last_error()
    el = np.array([6], np.uint16)
    assert not lib.init_d3_handle(d3_lib.ptr(el, C.c_uint16), 1, 10, -1.0, 40.0, 0, 0)
    assert not lib.init_d3_handle(d3_lib.ptr(el, C.c_uint16), 1, 10, 40.0, 40.0, 7, 0)
    assert not lib.init_d3_handle(d3_lib.ptr(el, C.c_uint16), 1, 10, 40.0, 40.0, 0, 42)
    assert not lib.d3x_init(d3_lib.ptr(el, C.c_uint16), 1, 10, 40.0, 40.0, 0.0, 0, 0, 5)


def test_null_handle_calls_fail_cleanly():
    lib = d3_lib.load()
    e = C.c_float()
    f = np.zeros(3, np.float32)
    s = np.zeros(9, np.float32)
    assert lib.compute_dispersion_energy_from_handle_status(None, C.byref(e), d3_lib.ptr(f, C.c_float),
                                                            d3_lib.ptr(s, C.c_float)) == 1
    lib.set_atoms(None, d3_lib.ptr(f, C.c_float), None, 0)
    lib.clear_d3_handle(None)
    lib.free_d3_handle(None)
    assert lib.d3x_compute_f64(None, None, None, None) != 0
    assert lib.d3x_last_status(None) != 0


def test_synthetic_table_invariants():
    t = d3_lib.table()
    assert t["name"] == "synthetic-original" and (t["nel"], t["nref"], t["zoff"]) == (95, 5, 0)
    c = t["c6ab"]
    zs = np.arange(1, t["zmax"] + 1)
    sub = c[1:, 1:]
    # C6 symmetry c6[A][B][a][b] = c6[B][A][b][a]
    np.testing.assert_array_equal(sub[..., 0], np.transpose(sub[..., 0], (1, 0, 3, 2)))
    # CN references depend only on (element, reference); -1 sentinels are consistent
    cnref = c[zs, zs, :, 0, 1]
    for A in zs:
        for B in zs[::7]:
            valid = (c[A, B, :, :, 1] > -1) & (c[A, B, :, :, 2] > -1)
            exp_valid = (cnref[A - 1][:, None] > -1) & (cnref[B - 1][None, :] > -1)
            np.testing.assert_array_equal(valid, exp_valid)
            np.testing.assert_array_equal(np.where(valid, c[A, B, :, :, 1], 0),
                                          np.where(valid, np.broadcast_to(cnref[A - 1][:, None], (5, 5)), 0))
            np.testing.assert_array_equal(c[A, B, :, :, 0][~valid], -1.0)
    # every element has its first reference at CN = 0; magnitudes in plausible ranges
    assert np.all(cnref[:, 0] == 0.0)
    v = sub[..., 0][sub[..., 1] > -1]
    assert v.min() >= 1.0 and v.max() <= 2000.0
    np.testing.assert_array_equal(t["r0ab"], t["r0ab"].T)
    assert 2.0 <= t["r0ab"][1:, 1:].min() and t["r0ab"][1:, 1:].max() <= 9.0
    assert 0.8 <= t["rcov"][1:].min() and t["rcov"][1:].max() <= 5.0
    assert 1.5 <= t["r2r4"][1:].min() and t["r2r4"][1:].max() <= 9.0


def test_table_file_round_trip(tmp_path):
    import pack_table
    t0 = d3_lib.table()
    path = str(tmp_path / "t.d3tab")
    pack_table.write_table(path, t0["c6ab"] * np.where(t0["c6ab"] > 0, 1.5, 1.0).astype(np.float32),
                           t0["r0ab"], t0["rcov"], t0["r2r4"], zoff=0, zmax=94)
    try:
        d3_lib.load_table(path)
        t1 = d3_lib.table()
        assert t1["name"] == path
        np.testing.assert_array_equal(t1["r0ab"], t0["r0ab"])
        assert np.allclose(t1["c6ab"][6, 6, 0, 0, 0], 1.5 * t0["c6ab"][6, 6, 0, 0, 0])
        d3_lib.load_table("synthetic-extended")
        te = d3_lib.table()
        assert (te["nel"], te["nref"], te["zoff"], te["zmax"]) == (103, 7, 1, 103)
        with pytest.raises(RuntimeError):
            d3_lib.load_table(str(tmp_path / "missing.d3tab"))
    finally:
        d3_lib.load_table(None)
    assert d3_lib.table()["name"] == "synthetic-original"


def test_non_factorising_table_passes_the_table_checks(tmp_path):
    """A table whose CN references are not one set per element (entry [1] of c6ab[A][B][a][b]
    differing between partners B) switches the handle to per-pair C6 interpolation instead of
    being refused: here, without a device, init gets past every table check and fails only at
    the first device call.  (The GPU tests evaluate such tables against the oracle.)"""
    import pack_table
    t0 = d3_lib.table()
    c6 = t0["c6ab"].copy()
    c6[6, 8, 0, 0, 1] += 0.25  # C's first CN reference, seen from a C-O pair only
    path = str(tmp_path / "nonfact.d3tab")
    pack_table.write_table(path, c6, t0["r0ab"], t0["rcov"], t0["r2r4"], zoff=0, zmax=94)
    lib = d3_lib.load()
    try:
        d3_lib.load_table(path)
        el = np.array([6, 8], np.uint16)
        h = lib.d3x_init(d3_lib.ptr(el, C.c_uint16), 2, 16, 40.0, 40.0, 0.0, 1, 0, 0)
        if not h:  # no device here
            err = d3_lib.last_error()
            assert "factorise" not in err and "table" not in err, err
        else:
            lib.free_d3_handle(h)
    finally:
        d3_lib.load_table(None)


def test_header_initialiser_parsing(tmp_path):
    """pack_table reads C initialisers of c6ab_ref / r0ab / rcov / r2r4 in declaration order."""
    import pack_table
    t = d3_lib.table()

    def arr(name, a, dims):
        body = ", ".join(f"{v:.9g}f" for v in np.asarray(a, np.float64).ravel())
        return f"static const float {name}{dims} = {{ {body} }}; // generated\n"

    hdr = tmp_path / "gen_constants_original.h"
    hdr.write_text("/* test header */\n" + arr("c6ab_ref", t["c6ab"], "[95][95][5][5][3]") +
                   arr("r0ab", t["r0ab"], "[95][95]") + arr("rcov", t["rcov"], "[95]") +
                   arr("r2r4", t["r2r4"], "[95]"))
    c6, r0, rc, rr = pack_table.read_header(str(hdr), extended=False)
    np.testing.assert_array_equal(c6.astype(np.float32), t["c6ab"])
    np.testing.assert_array_equal(r0.astype(np.float32), t["r0ab"])
    np.testing.assert_array_equal(rc.astype(np.float32), t["rcov"])
    np.testing.assert_array_equal(rr.astype(np.float32), t["r2r4"])
