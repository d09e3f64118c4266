"""The command-line driver on the GPU: the reference's example structure (192 atoms of TiO2,
tests/golden/tio2_192.extxyz) through lasp-d3_amd/build/d3_cpp with the reference driver's
settings (cutoff 60 bohr, CN and ATM 40 bohr, zero damping, PBE), against the FP64 oracle, on the
FP64 path (--fp64) and on the default FP32 path; the same structure written as .arc gives the same
energy, and a malformed .arc atom line is skipped without misaligning the others."""

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


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("atm", [False, True])
def test_cli_matches_oracle(table, atm, precision, tmp_path):
    """--fp64 at the float output's resolution; the default FP32 path within its 2e-6 / 2e-5 bar."""
    print(f"case cli tio2_192 atm={atm} {precision}", flush=True)
    (pos, Z, cell), = d3_io.read_extxyz(GOLDEN)
    args = (["--use-atm"] if atm else []) + (["--fp64"] if precision == "fp64" else [])
    te, tf, ts = (2e-7, 1e-5, 1e-6) if precision == "fp64" else (2e-6, 2e-5, 2e-5)
    e, f, s, out = run_cli(*args, GOLDEN)
    assert "Parsed 1 structures from file" in out and len(e) == 1 and f.shape == (192, 3)
    # the driver passes float coordinates and cell, as the reference's does
    pos32, cell32 = pos.astype(np.float32).astype(np.float64), cell.astype(np.float32).astype(np.float64)
    o = d3_oracle.compute(pos32, Z, cell32, table, cutoff=60.0, cn_cutoff=40.0, atm_cutoff=40.0 if atm else 0.0,
                          damping=0, functional=0)
    assert abs(e[0] - o["energy"]) <= te * abs(o["energy"])
    assert np.max(np.abs(f - o["forces"])) <= tf * max(1.0, np.max(np.abs(o["forces"])))
    assert np.max(np.abs(s.reshape(3, 3) - o["stress"])) <= ts * np.max(np.abs(o["stress"])) + 1e-9
    # the same structure as .arc (cell re-oriented): the same energy
    arc = str(tmp_path / "tio2.arc")
    d3_io.write_arc(arc, [(pos, Z, cell)])
    e2, _, _, _ = run_cli(*args, arc)
    assert abs(e2[0] - e[0]) <= 1e-6 * abs(e[0])


def test_cli_skips_malformed_arc_atom_line(table, tmp_path):
    """An .arc atom line whose coordinate is not a number is reported and skipped: the structure
    keeps its other atoms (element list and coordinates stay aligned) and evaluates like them."""
    print("case cli malformed arc line", flush=True)
    (pos, Z, cell), = d3_io.read_extxyz(GOLDEN)
    pos, Z = pos[:24], Z[:24]
    arc = tmp_path / "part.arc"
    d3_io.write_arc(str(arc), [(pos, Z, cell)])
    (p2, z2, c2), = d3_io.read_arc(str(arc))  # the structure as the .arc file holds it
    lines = arc.read_text().splitlines()
    k = next(i for i, ln in enumerate(lines) if ln.split()[:1] == ["PBC"])
    lines.insert(k + 1, "O     1.000000000    x.5000000    3.000000000 CORE    1 O  O   0.0000    1")
    arc.write_text("\n".join(lines) + "\n")
    e, f, _, out = run_cli("--fp64", str(arc))
    assert f.shape == (24, 3) and "for 24 atoms" in out
    p32, c32 = p2.astype(np.float32).astype(np.float64), c2.astype(np.float32).astype(np.float64)
    o = d3_oracle.compute(p32, z2, c32, table, cutoff=60.0, cn_cutoff=40.0, damping=0, functional=0)
    assert abs(e[0] - o["energy"]) <= 2e-7 * abs(o["energy"])
