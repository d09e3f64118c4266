"""Tuning parameters change only the summation order: results for several bin widths and
three-body grid sizes agree to the parity bar; the autotuner runs and its cache round-trips."""

import json
import os

import numpy as np
import pytest

import d3_lib
import systems

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [d3_lib.D3X_FP64, d3_lib.D3X_FP32])
def test_results_invariant_under_tuning(precision):
    s = systems.si64(seed=31)
    ref = None
    for w, b in ((None, None), (5.0, 1), (9.5, 7), (14.0, 64), (41.0, 4096), (60.0, 300)):
        h = d3_lib.Handle(s["Z"], len(s["Z"]), cutoff=s["cutoff"], cn_cutoff=s["cn_cutoff"], atm_cutoff=s["atm"],
                          damping=1, functional=0, precision=precision)
        if w is not None:
            h.set_tuning(w, b)
        h.set_atoms(s["pos"], s["Z"])
        h.set_cell(s["cell"])
        out = h.compute()
        h.close()
        if ref is None:
            ref = out
            continue
        # FP64: the parity bar; FP32: its own error budget (forces in this crystal nearly cancel,
        # so they get an absolute floor of 1e-6 eV/A, the size of FP32 force errors)
        fp64 = precision == d3_lib.D3X_FP64
        tol = 1e-12 if fp64 else 2e-6
        assert abs(out[0] - ref[0]) <= tol * abs(ref[0]), (w, b)
        assert np.max(np.abs(out[1] - ref[1])) <= (1e-9 if fp64 else 1e-6), (w, b)
        assert np.max(np.abs(out[2] - ref[2])) <= tol * 10 * np.max(np.abs(ref[2])) + (0 if fp64 else 1e-9), (w, b)


def test_autotune_and_cache(tmp_path):
    s = systems.triclinic_mixed(n=24, seed=8)
    cache = str(tmp_path / "tune.json")
    h = d3_lib.Handle(s["Z"], len(s["Z"]), cutoff=s["cutoff"], cn_cutoff=s["cn_cutoff"], atm_cutoff=s["atm"],
                      damping=1, functional=2)
    h.set_atoms(s["pos"], s["Z"])
    h.set_cell(s["cell"])
    e0, f0, _ = h.compute()
    t = h.autotune(cache)
    w, b = h.tuning()
    assert t > 0 and w > 0 and b > 0
    e1, f1, _ = h.compute()
    assert abs(e1 - e0) <= 1e-12 * abs(e0)
    h.close()
    data = json.load(open(cache))
    assert len(data) == 1
    (key, entry), = data.items()
    assert entry["bin_width"] == pytest.approx(w) and entry["atm_blocks"] == b
    assert entry["atm_variant"] in (0, 1, 2, 3)
    # a new handle with the cache in the environment picks the tuned values for the same input
    os.environ["D3_TUNE_CACHE"] = cache
    try:
        h2 = d3_lib.Handle(s["Z"], len(s["Z"]), cutoff=s["cutoff"], cn_cutoff=s["cn_cutoff"],
                           atm_cutoff=s["atm"], damping=1, functional=2)
        h2.set_atoms(s["pos"], s["Z"])
        h2.set_cell(s["cell"])
        e2, _, _ = h2.compute()
        assert h2.tuning() == (pytest.approx(w), b)
        assert h2.atm_variant() == entry["atm_variant"]
        assert abs(e2 - e0) <= 1e-12 * abs(e0)
        h2.close()
    finally:
        del os.environ["D3_TUNE_CACHE"]


def write_cache(path, data):
    """The engine's cache layout: a JSON object with one entry per line."""
    rows = [f"  {json.dumps(k)}: {json.dumps(v)}" for k, v in data.items()]
    with open(path, "w") as f:
        f.write("{\n" + ",\n".join(rows) + "\n}\n")


def test_every_three_body_variant(tmp_path):
    """Each three-body kernel variant (k-side sums in LDS or not, damping rows or not), forced
    through the tuning cache, gives the same result: FP64 at the parity bar, FP32 within its error
    budget of the FP64 result, and every FP32 variant the same FP32 tile sums as the first (the
    results then differ only by the order of FP64 sums, ~1e-16)."""
    s = systems.triclinic_mixed(n=24, seed=9)
    kw = dict(cutoff=s["cutoff"], cn_cutoff=s["cn_cutoff"], atm_cutoff=s["atm"], damping=1, functional=2)
    cache = str(tmp_path / "tune.json")
    ref = None
    for prec in (d3_lib.D3X_FP64, d3_lib.D3X_FP32):
        h = d3_lib.Handle(s["Z"], len(s["Z"]), precision=prec, **kw)
        h.set_atoms(s["pos"], s["Z"])
        h.set_cell(s["cell"])
        h.compute()
        h.autotune(cache)  # writes the entry (and its key) for this input and precision
        h.close()
    data = json.load(open(cache))
    assert len(data) == 2
    os.environ["D3_TUNE_CACHE"] = cache
    ref32 = None
    try:
        for prec in (d3_lib.D3X_FP64, d3_lib.D3X_FP32):
            for v in range(4):
                for key in data:
                    data[key]["atm_variant"] = v
                write_cache(cache, data)
                h = d3_lib.Handle(s["Z"], len(s["Z"]), precision=prec, **kw)
                h.set_atoms(s["pos"], s["Z"])
                h.set_cell(s["cell"])
                e, f, st = h.compute()
                assert h.atm_variant() == v
                h.close()
                if ref is None:
                    ref = (e, f, st)
                    continue
                fp64 = prec == d3_lib.D3X_FP64
                assert abs(e - ref[0]) <= (1e-12 if fp64 else 2e-6) * abs(ref[0]), (prec, v)
                assert np.max(np.abs(f - ref[1])) <= (1e-9 if fp64 else 2e-5 * np.max(np.abs(ref[1]))), (prec, v)
                if not fp64:
                    if ref32 is None:
                        ref32 = (e, f, st)
                    assert abs(e - ref32[0]) <= 1e-13 * abs(ref32[0]), v
                    assert np.max(np.abs(f - ref32[1])) <= 1e-12 * np.max(np.abs(ref32[1])), v
                    assert np.max(np.abs(st - ref32[2])) <= 1e-12 * np.max(np.abs(ref32[2])), v
    finally:
        del os.environ["D3_TUNE_CACHE"]


def test_default_bin_width_follows_cn_cutoff():
    """Default bins follow the CN cutoff alone, also when the two-body cutoff is smaller (a
    1-bohr two-body cutoff once shrank them to 4 bohr, and the CN, chain-rule and feature
    passes then walked thousands of near-empty bins per centre): half of it before any cell is
    known and on cells at least 6 CN cutoffs across, 3/8 of it on smaller cells such as this one."""
    s = systems.si64(seed=3)
    for rc2, rccn in ((94.87, 40.0), (1.0, 40.0), (50.0, 30.0)):
        h = d3_lib.Handle(s["Z"], len(s["Z"]), cutoff=rc2, cn_cutoff=rccn, atm_cutoff=0.0, damping=1, functional=0)
        w, _ = h.tuning()
        assert w == pytest.approx(0.5 * rccn), (rc2, rccn, w)
        h.set_atoms(s["pos"], s["Z"])
        h.set_cell(s["cell"])
        h.compute()
        w, _ = h.tuning()
        h.close()
        assert w == pytest.approx(0.375 * rccn), (rc2, rccn, w)


@pytest.mark.parametrize("system", ["si64", "water_small"])
def test_fp32_results_do_not_depend_on_work_items(system):
    """The three-body kernel's work items per centre follow the number of persistent blocks; their
    k-tile ranges start on even tiles, so the two-tile passes, and every FP32 tile sum, are the
    same for any number of them: FP32 results agree to the order of FP64 sums."""
    s = systems.si64(seed=31) if system == "si64" else systems.water_small(n_mol=32, seed=3)
    ref = None
    for b in (0, 64, 300, 4096):
        h = d3_lib.Handle(np.unique(s["Z"]), len(s["Z"]), cutoff=60.0, cn_cutoff=30.0, atm_cutoff=30.0, damping=1,
                          functional=0, precision=d3_lib.D3X_FP32)
        if b:
            h.set_tuning(0.0, b)
        h.set_atoms(s["pos"], s["Z"])
        h.set_cell(s["cell"])
        out = h.compute()
        h.close()
        if ref is None:
            ref = out
            continue
        assert abs(out[0] - ref[0]) <= 1e-13 * abs(ref[0]), b
        assert np.max(np.abs(out[1] - ref[1])) <= 1e-12 * np.max(np.abs(ref[1])), b
