"""Pins of the FP64 oracle (CPU): closed forms, finite differences and invariants, chosen so that
a dropped term, a wrong sign or index, or a transposed operand fails at least one of them.

Closed forms use single-reference elements (C: Z=6, B: Z=5 in the synthetic table), for which the
C6 interpolation returns the reference C6 exactly and dC6/dCN = 0, and a box so large that only
the direct pair lies within the cutoffs.  The mixed-element closed forms (an O-H-H cluster and an
O-N-H triangle, multi-reference elements with three different CNs) write out every CN, the
25-term C6 sums and their CN derivatives, so that an interpolation pairing an atom's CN with the
other element's references, or a three-body C6_jk taken at swapped CNs, fails.
"""

import numpy as np
import pytest

import d3_oracle
import systems

BOHR = 0.52917726
HARTREE = 27.211396641308
PBE0 = d3_oracle.functional_params(1)
PBE = d3_oracle.functional_params(0)


def run(pos, Z, cell, table, **kw):
    return d3_oracle.compute(pos, Z, cell, table, **kw)


def iso(pos, L=120.0):
    """Atoms placed in the middle of a cubic box of edge L Angstrom."""
    return np.asarray(pos, float) + L / 2, np.eye(3) * L


# ------------------------------------------------------------------------------- closed forms

@pytest.mark.parametrize("damping", [0, 1])
def test_dimer_energy_closed_form(table, damping):
    r = 2.7  # Angstrom
    pos, cell = iso([[0, 0, 0], [r, 0, 0]])
    Z = [6, 5]
    o = run(pos, Z, cell, table, cutoff=40.0, cn_cutoff=40.0, damping=damping, functional=1)
    c6 = float(table["c6ab"][6, 5, 0, 0, 0])
    q = float(table["r2r4"][6]) * float(table["r2r4"][5])
    c8 = 3.0 * c6 * q
    rb = r / BOHR
    s6, s8z, s8b, sr6, sr8, a1, a2 = PBE0
    if damping == 0:
        R0 = float(table["r0ab"][6, 5])
        f6 = 1 / (1 + 6 * (sr6 * R0 / rb) ** 14)
        f8 = 1 / (1 + 6 * (sr8 * R0 / rb) ** 16)
        s8 = s8z
    else:
        a = a1 * np.sqrt(3 * q) + a2
        f6 = rb ** 6 / (rb ** 6 + a ** 6)
        f8 = rb ** 8 / (rb ** 8 + a ** 8)
        s8 = s8b
    E = -HARTREE * (s6 * c6 * f6 / rb ** 6 + s8 * c8 * f8 / rb ** 8)
    assert o["energy"] == pytest.approx(E, rel=1e-13)
    # both atoms take half of the pair energy
    np.testing.assert_allclose(o["atom_energy"], [E / 2, E / 2], rtol=1e-13)


def test_pair_cn_closed_form(table):
    r = 1.9
    pos, cell = iso([[0, 0, 0], [0, r, 0]])
    o = run(pos, [6, 8], cell, table, cutoff=40.0, cn_cutoff=40.0, damping=1)
    R = float(table["rcov"][6]) + float(table["rcov"][8])
    cn = 1.0 / (1.0 + np.exp(-16.0 * (R / (r / BOHR) - 1.0)))
    np.testing.assert_allclose(o["cn"], [cn, cn], rtol=1e-14)


def test_smooth_cn_cutoff_closed_form(table):
    """The reference's optional smooth CN cutoff (SMOOTH_CUTOFF_CN, src/d3_kernel.cu:550-561):
    an isolated pair's CN is tanh^3(rc_CN - r) / (1 + exp(-k1 (R / r - 1))) with r, rc_CN in bohr,
    and a pair just inside the cutoff contributes next to nothing."""
    r = 1.9
    pos, cell = iso([[0, 0, 0], [0, r, 0]])
    rc = 6.0  # bohr
    o = run(pos, [6, 8], cell, table, cutoff=40.0, cn_cutoff=rc, damping=1, smooth_cn=True)
    R = float(table["rcov"][6]) + float(table["rcov"][8])
    rb = r / BOHR
    cn = np.tanh(rc - rb) ** 3 / (1.0 + np.exp(-16.0 * (R / rb - 1.0)))
    np.testing.assert_allclose(o["cn"], [cn, cn], rtol=1e-13)
    sharp = run(pos, [6, 8], cell, table, cutoff=40.0, cn_cutoff=rc, damping=1)
    assert abs(o["cn"][0] - sharp["cn"][0]) > 1e-3  # the option changes the result here
    edge = run(iso([[0, 0, 0], [0, (rc - 1e-4) * BOHR, 0]])[0], [6, 8], iso([[0, 0, 0]])[1], table, cutoff=40.0,
               cn_cutoff=rc, damping=1, smooth_cn=True)
    assert 0 < edge["cn"][0] < 1e-11


@pytest.mark.parametrize("atm", [False, True])
def test_smooth_cn_forces_and_stress_match_fd(table, atm):
    pos, Z, cell, kw = small_triclinic(atm)
    # a CN cutoff (7 bohr) where the terms near it are large enough for the smooth factor to count
    kw = dict(kw, cn_cutoff=7.0, smooth_cn=True)
    o = run(pos, Z, cell, table, **kw)
    h = 1e-4
    for i in (0, 3, 5):
        for c in range(3):
            p1, p2 = pos.copy(), pos.copy()
            p1[i, c] += h
            p2[i, c] -= h
            fd = -(run(p1, Z, cell, table, **kw)["energy"] - run(p2, Z, cell, table, **kw)["energy"]) / (2 * h)
            assert abs(fd - o["forces"][i, c]) < 1e-6, (i, c)
    V = abs(np.linalg.det(cell))
    h = 1e-5
    for a, b in ((0, 0), (1, 2)):
        eps = np.zeros((3, 3))
        eps[a, b] = h
        ep = run(pos @ (np.eye(3) + eps), Z, cell @ (np.eye(3) + eps), table, **kw)["energy"]
        em = run(pos @ (np.eye(3) - eps), Z, cell @ (np.eye(3) - eps), table, **kw)["energy"]
        assert abs((ep - em) / (2 * h) / V - o["stress"][a, b]) < 1e-8 * max(1.0, np.max(np.abs(o["stress"]))) + 1e-9


def test_cn_cutoff_is_inclusive_and_twobody_cutoff_strict(table):
    """CN counts r^2 <= rc^2, the two-body sum r^2 < rc^2 (src/d3_kernel.cu:626, :1073)."""
    rb = 12.0
    pos, cell = iso([[0, 0, 0], [rb * BOHR, 0, 0]])
    inside = run(pos, [6, 6], cell, table, cutoff=rb * 1.000001, cn_cutoff=rb * 1.000001, damping=1)
    edge2b = run(pos, [6, 6], cell, table, cutoff=rb * 0.999999, cn_cutoff=rb * 1.000001, damping=1)
    assert inside["n2b"] == 2 and inside["ncn"] == 2
    assert edge2b["n2b"] == 0 and edge2b["energy"] == 0.0


def test_two_reference_c6_by_hand(table):
    """Si has two references: C6(CN_A, CN_B) = sum C6ref L / sum L with L = exp(-4 dCN^2 ...)."""
    r = 2.35
    pos, cell = iso([[0, 0, 0], [r, 0, 0]])
    o = run(pos, [14, 14], cell, table, cutoff=40.0, cn_cutoff=40.0, damping=0, functional=1)
    cn = o["cn"][0]
    c = table["c6ab"][14, 14]
    num = den = 0.0
    for a in range(2):
        for b in range(2):
            L = np.exp(-4.0 * ((cn - c[a, b, 1]) ** 2 + (cn - c[a, b, 2]) ** 2))
            num += c[a, b, 0] * L
            den += L
    c6 = num / den
    rb = r / BOHR
    s6, s8z, _, sr6, sr8, _, _ = PBE0
    R0 = float(table["r0ab"][14, 14])
    q = float(table["r2r4"][14]) ** 2
    f6 = 1 / (1 + 6 * (sr6 * R0 / rb) ** 14)
    f8 = 1 / (1 + 6 * (sr8 * R0 / rb) ** 16)
    E = -HARTREE * (s6 * c6 * f6 / rb ** 6 + s8z * 3 * c6 * q * f8 / rb ** 8)
    assert o["energy"] == pytest.approx(E, rel=1e-12)


def test_bj_short_range_limit(table):
    """BJ damping keeps the energy finite: E -> -(s6 C6 / a^6 + s8 C8 / a^8), a = a1 R0 + a2."""
    pos, cell = iso([[0, 0, 0], [1e-4, 0, 0]])
    o = run(pos, [6, 5], cell, table, cutoff=40.0, cn_cutoff=40.0, damping=1, functional=0)
    c6 = float(table["c6ab"][6, 5, 0, 0, 0])
    q = float(table["r2r4"][6]) * float(table["r2r4"][5])
    s6, _, s8, _, _, a1, a2 = PBE
    a = a1 * np.sqrt(3 * q) + a2
    lim = -HARTREE * (s6 * c6 / a ** 6 + s8 * 3 * c6 * q / a ** 8)
    assert o["energy"] == pytest.approx(lim, rel=1e-5)


def _atm_energy(table, pos):
    pos, cell = iso(pos)
    kw = dict(cutoff=40.0, cn_cutoff=40.0, damping=1, functional=0)
    with_atm = run(pos, [6, 6, 6], cell, table, atm_cutoff=40.0, **kw)
    without = run(pos, [6, 6, 6], cell, table, **kw)
    return with_atm["energy"] - without["energy"], with_atm


@pytest.mark.parametrize("geom", ["equilateral", "collinear"])
def test_atm_closed_form(table, geom):
    s = 3.1
    if geom == "equilateral":
        pos = [[0, 0, 0], [s, 0, 0], [s / 2, s * np.sqrt(3) / 2, 0]]
        edges = np.array([s, s, s]) / BOHR
        cosprod = 1.0 / 8.0
    else:
        pos = [[0, 0, 0], [s, 0, 0], [2 * s, 0, 0]]
        edges = np.array([s, s, 2 * s]) / BOHR
        cosprod = -1.0
    dE, o = _atm_energy(table, pos)
    c6 = float(table["c6ab"][6, 6, 0, 0, 0])
    C9 = -c6 ** 1.5
    A = (1 + 3 * cosprod) / np.prod(edges) ** 3
    R0 = float(table["r0ab"][6, 6])
    f = 1 / (1 + 6 * (4.0 / 3.0 * R0 / np.prod(edges) ** (1 / 3)) ** 16)
    assert dE == pytest.approx(-HARTREE * C9 * A * f, rel=1e-12)
    assert o["ntri"] == 3  # one triangle, counted once per vertex


# Mixed-element closed forms.  Every quantity below is written out from the definitions in
# src/d3_kernel.cu (CN :544-574, C6 :709-800, two-body :902-986, ATM :1199-1509) with each
# element's CN references read from its homonuclear diagonal entry c6ab[Z][Z][a][a][1], so a
# C6 interpolation that pairs CN_A with element B's references (or takes C6_jk at swapped CNs)
# fails: the clusters are shaped so that every atom has a different CN.

def _cn_refs(table, Z):
    """CN references of element Z (one per valid reference) from the homonuclear entries."""
    c = table["c6ab"]
    return [(a, float(c[Z, Z, a, a, 1])) for a in range(table["nref"]) if c[Z, Z, a, a, 1] > -1]


def _pair_cn(table, Zi, Zj, r_bohr):
    R = float(table["rcov"][Zi]) + float(table["rcov"][Zj])
    return 1.0 / (1.0 + np.exp(-16.0 * (R / r_bohr - 1.0)))


def _c6_closed(table, Zi, Zj, cni, cnj):
    """C6(CN_i, CN_j) = sum_ab C6ref_ab L_ab / sum_ab L_ab with
    L_ab = exp(-4 (CN_i - CNref_i,a)^2 - 4 (CN_j - CNref_j,b)^2), and dC6/dCN_i, dC6/dCN_j."""
    num = den = dnum_i = dden_i = dnum_j = dden_j = 0.0
    for a, ra in _cn_refs(table, Zi):
        for b, rb in _cn_refs(table, Zj):
            L = np.exp(-4.0 * (cni - ra) ** 2 - 4.0 * (cnj - rb) ** 2)
            cref = float(table["c6ab"][Zi, Zj, a, b, 0])
            num += cref * L
            den += L
            dnum_i += cref * (-8.0 * (cni - ra)) * L
            dden_i += -8.0 * (cni - ra) * L
            dnum_j += cref * (-8.0 * (cnj - rb)) * L
            dden_j += -8.0 * (cnj - rb) * L
    c6 = num / den
    return c6, (dnum_i - c6 * dden_i) / den, (dnum_j - c6 * dden_j) / den


def test_mixed_element_table_invariants(table):
    """The entries [1] and [2] of c6ab[A][B][a][b] are CNref_A,a and CNref_B,b, and C6ref is
    symmetric under (A, a) <-> (B, b): the layout the closed forms below rely on."""
    c = table["c6ab"]
    for A, B in [(8, 1), (1, 8), (7, 8), (1, 7)]:
        ra, rb = dict(_cn_refs(table, A)), dict(_cn_refs(table, B))
        for a in ra:
            for b in rb:
                assert c[A, B, a, b, 1] == ra[a] and c[A, B, a, b, 2] == rb[b]
                assert c[A, B, a, b, 0] == c[B, A, b, a, 0]


@pytest.mark.parametrize("damping", [0, 1])
def test_mixed_element_cluster_closed_form(table, damping):
    """Isolated O-H-H cluster (two multi-reference elements, CN_O != CN_H1 != CN_H2): CN, C6 of
    each pair by the 25-term sum, energy, per-atom energy and dE/dCN written out."""
    pos = np.array([[0.0, 0.0, 0.0], [1.60, 0.0, 0.0], [-0.55, 1.78, 0.30]])
    Z = [8, 1, 1]
    P, cell = iso(pos)
    o = run(P, Z, cell, table, cutoff=40.0, cn_cutoff=40.0, damping=damping, functional=0)
    s6, s8z, s8b, sr6, sr8, a1, a2 = PBE
    n = 3
    rb = {(i, j): np.linalg.norm(pos[i] - pos[j]) / BOHR for i in range(n) for j in range(n) if i != j}
    cn = np.array([sum(_pair_cn(table, Z[i], Z[j], rb[i, j]) for j in range(n) if j != i) for i in range(n)])
    np.testing.assert_allclose(o["cn"], cn, rtol=1e-12)
    assert min(abs(cn[0] - cn[1]), abs(cn[0] - cn[2]), abs(cn[1] - cn[2])) > 0.1
    E = 0.0
    ea = np.zeros(n)
    g = np.zeros(n)
    for i in range(n):
        for j in range(i + 1, n):
            r = rb[i, j]
            c6, dc6i, dc6j = _c6_closed(table, Z[i], Z[j], cn[i], cn[j])
            q = float(table["r2r4"][Z[i]]) * float(table["r2r4"][Z[j]])
            if damping == 0:
                R0 = float(table["r0ab"][Z[i], Z[j]])
                f6 = 1 / (1 + 6 * (sr6 * R0 / r) ** 14)
                f8 = 1 / (1 + 6 * (sr8 * R0 / r) ** 16)
                s8 = s8z
            else:
                a = a1 * np.sqrt(3 * q) + a2
                f6 = r ** 6 / (r ** 6 + a ** 6)
                f8 = r ** 8 / (r ** 8 + a ** 8)
                s8 = s8b
            t = s6 * f6 / r ** 6 + s8 * 3 * q * f8 / r ** 8  # E+ of the pair per unit C6
            E += c6 * t
            ea[i] += 0.5 * c6 * t
            ea[j] += 0.5 * c6 * t
            g[i] += dc6i * t
            g[j] += dc6j * t
    assert o["energy"] == pytest.approx(-HARTREE * E, rel=1e-13)
    np.testing.assert_allclose(o["atom_energy"], -HARTREE * ea, rtol=1e-13)
    np.testing.assert_allclose(o["dedcn"], g, rtol=1e-11, atol=1e-16)


def test_mixed_element_c6_derivative_closed_form_by_differences(table):
    """The closed form's dC6/dCN against central differences of its own C6 (so the dE/dCN pin
    above does not rest on an unchecked derivative)."""
    for Zi, Zj, ci, cj in [(8, 1, 0.83, 0.41), (7, 8, 1.9, 2.7), (1, 7, 0.2, 1.3)]:
        c6, di, dj = _c6_closed(table, Zi, Zj, ci, cj)
        h = 1e-6
        fi = (_c6_closed(table, Zi, Zj, ci + h, cj)[0] - _c6_closed(table, Zi, Zj, ci - h, cj)[0]) / (2 * h)
        fj = (_c6_closed(table, Zi, Zj, ci, cj + h)[0] - _c6_closed(table, Zi, Zj, ci, cj - h)[0]) / (2 * h)
        assert di == pytest.approx(fi, rel=1e-7, abs=1e-9 * abs(c6))
        assert dj == pytest.approx(fj, rel=1e-7, abs=1e-9 * abs(c6))


def test_mixed_element_atm_closed_form(table):
    """Isolated O-N-H triangle (three multi-reference elements, three different CNs): the ATM
    energy, its per-atom thirds and its dE/dCN from C9 = -sqrt(C6_ab C6_ac C6_bc) at the CNs,
    the angle factor from the triangle's cosines and the damping from r0ab."""
    pos = np.array([[0.0, 0.0, 0.0], [2.05, 0.0, 0.0], [-0.5, 1.2, 0.3]])
    Z = [8, 7, 1]
    P, cell = iso(pos)
    kw = dict(cutoff=40.0, cn_cutoff=40.0, damping=1, functional=0)
    with_atm = run(P, Z, cell, table, atm_cutoff=40.0, **kw)
    without = run(P, Z, cell, table, **kw)
    n = 3
    r = {(i, j): np.linalg.norm(pos[i] - pos[j]) / BOHR for i in range(n) for j in range(n) if i != j}
    cn = np.array([sum(_pair_cn(table, Z[i], Z[j], r[i, j]) for j in range(n) if j != i) for i in range(n)])
    np.testing.assert_allclose(with_atm["cn"], cn, rtol=1e-12)
    assert min(abs(cn[0] - cn[1]), abs(cn[0] - cn[2]), abs(cn[1] - cn[2])) > 0.1
    pairs = [(0, 1), (0, 2), (1, 2)]
    c6 = {}
    for i, j in pairs:
        c, di, dj = _c6_closed(table, Z[i], Z[j], cn[i], cn[j])
        c6[i, j] = (c, {i: di, j: dj})
    C9 = -np.sqrt(c6[0, 1][0] * c6[0, 2][0] * c6[1, 2][0])
    # cosines of the three interior angles from the position vectors
    cos = []
    for v in range(3):
        u, w = [pos[x] - pos[v] for x in range(3) if x != v]
        cos.append(np.dot(u, w) / (np.linalg.norm(u) * np.linalg.norm(w)))
    rp = r[0, 1] * r[0, 2] * r[1, 2]
    A = (1 + 3 * cos[0] * cos[1] * cos[2]) / rp ** 3
    R0 = (float(table["r0ab"][Z[0], Z[1]]) * float(table["r0ab"][Z[0], Z[2]]) * float(table["r0ab"][Z[1], Z[2]])) ** (1 / 3)
    f = 1 / (1 + 6 * (4.0 / 3.0 * R0 / rp ** (1 / 3)) ** 16)
    e = C9 * A * f  # E+ of the triangle
    assert with_atm["energy"] - without["energy"] == pytest.approx(-HARTREE * e, rel=1e-11)
    np.testing.assert_allclose(with_atm["atom_energy"] - without["atom_energy"], [-HARTREE * e / 3] * 3, rtol=1e-10)
    # dE+/dCN_x = (C9 / 2) sum over the two edges at x of (dC6_xy/dCN_x) / C6_xy, times A f
    gx = np.zeros(n)
    for (i, j), (c, d) in c6.items():
        for x in (i, j):
            gx[x] += 0.5 * C9 * d[x] / c * A * f
    np.testing.assert_allclose(with_atm["dedcn"] - without["dedcn"], gx, rtol=1e-9, atol=1e-18)


# ---------------------------------------------------------------------- finite differences

def small_triclinic(atm):
    rng = np.random.Generator(np.random.PCG64(42))
    cell = np.array([[6.5, 0.0, 0.0], [1.2, 6.0, 0.0], [-0.8, 0.9, 5.8]])
    import d3_structures as S
    pos, _ = S.random_packing(7, cell, rng, d_min=1.6)
    Z = np.array([6, 8, 1, 14, 29, 7, 1], np.uint16)
    kw = dict(cutoff=30.0, cn_cutoff=20.0, atm_cutoff=18.0 if atm else 0.0, damping=1, functional=2)
    return pos, Z, cell, kw


@pytest.mark.parametrize("atm", [False, True])
def test_forces_match_finite_differences(table, atm):
    pos, Z, cell, kw = small_triclinic(atm)
    o = run(pos, Z, cell, table, **kw)
    h = 1e-4
    worst = 0.0
    for i in range(len(Z)):
        for c in range(3):
            p1, p2 = pos.copy(), pos.copy()
            p1[i, c] += h
            p2[i, c] -= h
            fd = -(run(p1, Z, cell, table, **kw)["energy"] - run(p2, Z, cell, table, **kw)["energy"]) / (2 * h)
            worst = max(worst, abs(fd - o["forces"][i, c]))
    assert worst < 1e-6, worst


def test_atm_forces_with_self_images(table):
    """Cell narrower than the three-body cutoff: triangles that contain the same atom twice
    (index-degenerate) must give forces consistent with the energy (the reference's forces are
    off by the degeneracy factor there)."""
    cell = np.array([[3.3, 0.0, 0.0], [0.4, 3.1, 0.0], [0.2, -0.3, 3.6]])
    pos = np.array([[0.3, 0.2, 0.1], [1.9, 1.4, 1.2]])
    Z = np.array([6, 14], np.uint16)
    kw = dict(cutoff=20.0, cn_cutoff=15.0, atm_cutoff=12.0, damping=0, functional=0)
    o = run(pos, Z, cell, table, **kw)
    h = 1e-4
    for i in range(2):
        for c in range(3):
            p1, p2 = pos.copy(), pos.copy()
            p1[i, c] += h
            p2[i, c] -= h
            fd = -(run(p1, Z, cell, table, **kw)["energy"] - run(p2, Z, cell, table, **kw)["energy"]) / (2 * h)
            assert abs(fd - o["forces"][i, c]) < 1e-6 * max(1.0, abs(fd))


@pytest.mark.parametrize("atm", [False, True])
def test_stress_matches_strain_derivative(table, atm):
    """sigma = (1/V) dE/d(strain) with positions and cell strained together."""
    pos, Z, cell, kw = small_triclinic(atm)
    o = run(pos, Z, cell, table, **kw)
    V = abs(np.linalg.det(cell))
    h = 1e-5
    for a in range(3):
        for b in range(3):
            eps = np.zeros((3, 3))
            eps[a, b] = h
            Fp, Fm = np.eye(3) + eps, np.eye(3) - eps
            ep = run(pos @ Fp, Z, cell @ Fp, table, **kw)["energy"]
            em = run(pos @ Fm, Z, cell @ Fm, table, **kw)["energy"]
            fd = (ep - em) / (2 * h) / V
            assert abs(fd - o["stress"][a, b]) < 1e-8 * max(1.0, np.max(np.abs(o["stress"]))) + 1e-9


# ------------------------------------------------------------------------------- invariants

def test_translation_permutation_rotation(table):
    pos, Z, cell, kw = small_triclinic(True)
    o = run(pos, Z, cell, table, **kw)
    # translation by a non-lattice vector
    t = run(pos + np.array([0.37, -1.91, 4.22]), Z, cell, table, **kw)
    assert t["energy"] == pytest.approx(o["energy"], rel=1e-12)
    np.testing.assert_allclose(t["forces"], o["forces"], atol=1e-11)
    # permutation
    p = np.random.Generator(np.random.PCG64(3)).permutation(len(Z))
    q = run(pos[p], Z[p], cell, table, **kw)
    assert q["energy"] == pytest.approx(o["energy"], rel=1e-12)
    np.testing.assert_allclose(q["forces"], o["forces"][p], atol=1e-11)
    # rigid rotation of atoms and cell together (row vectors: r' = r R^T)
    th = 0.7
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]]) @ \
        np.array([[1, 0, 0], [0, np.cos(0.3), -np.sin(0.3)], [0, np.sin(0.3), np.cos(0.3)]])
    r = run(pos @ R.T, Z, cell @ R.T, table, **kw)
    assert r["energy"] == pytest.approx(o["energy"], rel=1e-12)
    np.testing.assert_allclose(r["forces"], o["forces"] @ R.T, atol=1e-10)
    np.testing.assert_allclose(r["stress"], R @ o["stress"] @ R.T, atol=1e-11)


def test_supercell_extensive(table):
    pos, Z, cell, kw = small_triclinic(True)
    o = run(pos, Z, cell, table, **kw)
    sp, sz = [], []
    for x in range(2):
        for y in range(2):
            for z in range(2):
                sp.append(pos + x * cell[0] + y * cell[1] + z * cell[2])
                sz.append(Z)
    s = run(np.concatenate(sp), np.concatenate(sz), cell * 2, table, **kw)
    assert s["energy"] == pytest.approx(8 * o["energy"], rel=1e-11)
    np.testing.assert_allclose(s["forces"], np.tile(o["forces"], (8, 1)), atol=1e-10)
    np.testing.assert_allclose(s["stress"], o["stress"], atol=1e-11)


def test_per_centre_entry_equals_full_evaluation(table):
    """d3o_centers (per-centre passes for checks of large cells) gives, with the full evaluation's
    CN and dE/dCN as the other atoms' values, exactly the full evaluation's per-atom CN, energy,
    dE/dCN and force at the chosen centres."""
    pos, Z, cell, kw = small_triclinic(True)
    o = run(pos, Z, cell, table, **kw)
    cen = np.array([0, 2, 5, 6])
    c = d3_oracle.centers(pos, Z, cell, table, cen, o["cn"], o["dedcn"], **kw)
    np.testing.assert_allclose(c["cn"], o["cn"][cen], rtol=1e-14)
    np.testing.assert_allclose(c["atom_energy"], o["atom_energy"][cen], rtol=1e-13)
    np.testing.assert_allclose(c["dedcn"], o["dedcn"][cen], rtol=1e-12)
    np.testing.assert_allclose(c["forces"], o["forces"][cen], rtol=1e-12, atol=1e-15)


def test_left_handed_cell_flips_stress_sign(table):
    """V is the signed determinant (src/d3_kernel.cu:90-96): swapping two lattice vectors keeps
    the energy and forces and negates the stress (reference behaviour, quirk Q14)."""
    pos, Z, cell, kw = small_triclinic(False)
    o = run(pos, Z, cell, table, **kw)
    sw = run(pos, Z, cell[[1, 0, 2]], table, **kw)
    assert sw["energy"] == pytest.approx(o["energy"], rel=1e-12)
    np.testing.assert_allclose(sw["stress"], -o["stress"], atol=1e-12)


# ------------------------------------------- the reference's property tests (test/d3_test.cpp)

REF = systems.reference_systems()


@pytest.mark.parametrize("s", REF, ids=[s["name"] for s in REF])
def test_reference_properties(table, s):
    kw = dict(cutoff=s["cutoff"], cn_cutoff=s["cn_cutoff"], damping=s["damping"], functional=s["functional"])
    o = run(s["pos"], s["Z"], s["cell"], table, **kw)
    # EnergyNegative (:83-87)
    assert o["energy"] < 0
    # ResultStable (:89-129): bitwise repeatable
    o2 = run(s["pos"], s["Z"], s["cell"], table, **kw)
    assert o2["energy"] == o["energy"] and np.array_equal(o2["forces"], o["forces"])
    # ForceSumToZero (:131-147), StressSymmetry (:149-161), tightened for FP64
    assert np.max(np.abs(o["forces"].sum(axis=0))) < 1e-10
    np.testing.assert_allclose(o["stress"], o["stress"].T, atol=1e-13)
    # NumericForceMatch (:163-226): central differences, delta 1e-3 A, tolerance 1e-3 eV/A
    d = 1e-3
    for i in range(len(s["Z"])):
        for c in range(3):
            p1, p2 = s["pos"].copy(), s["pos"].copy()
            p1[i, c] += d
            p2[i, c] -= d
            fd = -(run(p1, s["Z"], s["cell"], table, **kw)["energy"] -
                   run(p2, s["Z"], s["cell"], table, **kw)["energy"]) / (2 * d)
            assert abs(fd - o["forces"][i, c]) < 1e-3
    # SupercellConsistency (:228-341): 2x2x2 (the reference allows 20 % on the energy ratio)
    sp, sz = [], []
    for x in range(2):
        for y in range(2):
            for z in range(2):
                sp.append(s["pos"] + x * s["cell"][0] + y * s["cell"][1] + z * s["cell"][2])
                sz.append(s["Z"])
    big = run(np.concatenate(sp), np.concatenate(sz), s["cell"] * 2, table, **kw)
    assert big["energy"] / o["energy"] == pytest.approx(8.0, rel=1e-10)
    np.testing.assert_allclose(big["forces"], np.tile(o["forces"], (8, 1)), atol=1e-9)
    np.testing.assert_allclose(big["stress"], o["stress"], atol=1e-12)
