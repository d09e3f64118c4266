"""The command-line driver on the GPU: the reference's example structure (192 atoms of TiO2,
tests/golden/tio2_192.extxyz) through lasp-d3_amd/build/d3_cpp with the reference driver's
settings (cutoff 60 bohr, CN and ATM 40 bohr, zero damping, PBE), against the FP64 oracle; the
same structure written as .arc gives the same energy."""

import os
import re
import subprocess

import numpy as np
import pytest

import d3_io
import d3_oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden", "tio2_192.extxyz")
CLI = os.path.join(ROOT, "lasp-d3_amd", "build", "d3_cpp")


def run_cli(*args):
    r = subprocess.run([CLI, *args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    out = r.stdout
    e = [float(x) for x in re.findall(r"^Energy: (\S+) eV", out, re.M)]
    f = np.array([[float(v) for v in m.split()] for m in re.findall(r"^Force\[\d+\]: (.*)$", out, re.M)])
    s = np.array([float(x) for x in re.findall(r"^Stress\[\d\]\[\d\]: (\S+)", out, re.M)])
    return e, f, s, out


@pytest.mark.parametrize("atm", [False, True])
def test_cli_matches_oracle(table, atm, tmp_path):
    print(f"case cli tio2_192 atm={atm}", flush=True)
    (pos, Z, cell), = d3_io.read_extxyz(GOLDEN)
    args = ["--use-atm"] if atm else []
    e, f, s, out = run_cli(*args, GOLDEN)
    assert "Parsed 1 structures from file" in out and len(e) == 1 and f.shape == (192, 3)
    # the driver passes float coordinates and cell, as the reference's does
    pos32, cell32 = pos.astype(np.float32).astype(np.float64), cell.astype(np.float32).astype(np.float64)
    o = d3_oracle.compute(pos32, Z, cell32, table, cutoff=60.0, cn_cutoff=40.0, atm_cutoff=40.0 if atm else 0.0,
                          damping=0, functional=0)
    assert abs(e[0] - o["energy"]) <= 2e-7 * abs(o["energy"])
    assert np.max(np.abs(f - o["forces"])) <= 1e-5 * max(1.0, np.max(np.abs(o["forces"])))
    assert np.max(np.abs(s.reshape(3, 3) - o["stress"])) <= 1e-6 * np.max(np.abs(o["stress"])) + 1e-9
    # the same structure as .arc (cell re-oriented): the same energy
    arc = str(tmp_path / "tio2.arc")
    d3_io.write_arc(arc, [(pos, Z, cell)])
    e2, _, _, _ = run_cli(*args, arc)
    assert abs(e2[0] - e[0]) <= 1e-6 * abs(e[0])
