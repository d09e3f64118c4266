"""Structure files and the command-line driver, on the CPU: the reference's example structure
(tests/golden/tio2_192.extxyz, a copy of the reference's examples/structures/192.extxyz), cell
parameters <-> lattice vectors, .arc round trip, and the driver's argument handling."""

import os
import subprocess

import numpy as np
import pytest

import d3_io
import d3_lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden", "tio2_192.extxyz")
CLI = os.environ.get("D3_CLI", os.path.join(ROOT, "lasp-d3_amd", "build", "d3_cpp"))


def test_reads_reference_example():
    (pos, Z, cell), = d3_io.read_extxyz(GOLDEN)
    assert pos.shape == (192, 3) and len(Z) == 192
    assert (Z == 8).sum() == 128 and (Z == 22).sum() == 64
    np.testing.assert_allclose(cell[0], [21.9552, 0.0, 0.0])
    np.testing.assert_allclose(cell[2], [2.569457538, -1.023176199, 6.296358256])
    np.testing.assert_allclose(pos[0], [7.683905, 5.285519, 2.769997])


def test_cell_parameters_round_trip():
    rng = np.random.default_rng(3)
    for _ in range(20):
        par = (*rng.uniform(3, 20, 3), *rng.uniform(60, 120, 3))
        cell = d3_io.cell_from_parameters(*par)
        np.testing.assert_allclose(d3_io.cell_parameters(cell), par, rtol=1e-10)
        # the volume of a parallelepiped from its parameters
        a, b, c, al, be, ga = par
        ca, cb, cg = np.cos(np.radians([al, be, ga]))
        v = a * b * c * np.sqrt(1 - ca * ca - cb * cb - cg * cg + 2 * ca * cb * cg)
        assert abs(np.linalg.det(cell) - v) <= 1e-9 * v
    np.testing.assert_allclose(d3_io.cell_from_parameters(4, 5, 6, 90, 90, 90), np.diag([4.0, 5.0, 6.0]), atol=1e-12)


def test_arc_round_trip(tmp_path):
    (pos, Z, cell), = d3_io.read_extxyz(GOLDEN)
    path = str(tmp_path / "tio2.arc")
    d3_io.write_arc(path, [(pos, Z, cell), (pos[:10], Z[:10], cell)])
    frames = d3_io.read_arc(path)
    assert len(frames) == 2 and len(frames[1][1]) == 10
    p2, z2, c2 = frames[0]
    np.testing.assert_array_equal(z2, Z)
    # same cell up to a rotation: same parameters, and the same fractional coordinates
    np.testing.assert_allclose(d3_io.cell_parameters(c2), d3_io.cell_parameters(cell), rtol=1e-9)
    np.testing.assert_allclose(p2 @ np.linalg.inv(c2), pos @ np.linalg.inv(cell), atol=1e-8)


def test_cli_arguments():
    if not os.path.exists(CLI):
        pytest.skip("d3_cpp not built")
    r = subprocess.run([CLI, "--help"], capture_output=True, text=True)
    assert r.returncode == 0 and "Usage" in r.stdout
    r = subprocess.run([CLI], capture_output=True, text=True)
    assert r.returncode != 0 and "No input file" in r.stderr
    r = subprocess.run([CLI, "/nonexistent.arc"], capture_output=True, text=True)
    assert r.returncode == 1 and "does not exist" in r.stderr


def test_cli_rejects_malformed_extxyz(tmp_path):
    """A coordinate that is not a number ends the parse with an error (no evaluation)."""
    if not os.path.exists(CLI):
        pytest.skip("d3_cpp not built")
    bad = tmp_path / "bad.extxyz"
    bad.write_text('2\nLattice="10 0 0 0 10 0 0 0 10" Properties=species:S:1:pos:R:3\nC 0 0 0\nO 1.0 abc 0\n')
    r = subprocess.run([CLI, str(bad)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "Invalid line" in r.stderr and "Error parsing file" in r.stderr


def test_cli_declares_only_the_reference_api():
    """The driver is a caller of the reference's C API only (include/d3.h), like the reference's."""
    src = open(os.path.join(ROOT, "lasp-d3_amd", "cli", "d3_cli.cpp")).read()
    assert '#include "../../include/d3.h"' in src and "d3x" not in src.replace("d3_cpp", "")
    assert d3_lib.load() is not None
