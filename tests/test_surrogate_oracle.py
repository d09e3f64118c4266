"""Pins of the CPU restatement of the learned three-body stand-in (oracle/surrogate_oracle.py):
its CN equals the D3 oracle's, its input and parameter gradients equal finite differences of its
forward pass, its forces and stress equal finite differences of its energy, and the features
have the closed-form values of isolated pairs.  CPU only."""

import numpy as np
import pytest

import d3_lib
import d3_oracle
import surrogate_oracle as SO
import systems


@pytest.fixture(scope="module")
def rcov(table):
    return np.asarray(table["rcov"], np.float64)


def small(seed=3, n=12):
    s = systems.triclinic_mixed(n=n, seed=seed)
    return s["pos"], s["Z"], s["cell"]


def random_params(seed=1, n=20):
    rng = np.random.default_rng(seed)
    p = SO.init_params(seed, mean=rng.normal(0, 1, SO.N_FEAT), istd=rng.uniform(0.5, 2, SO.N_FEAT),
                       scale=[0.01, 0.02, 0.03], shift=[0.001, -0.002, 0.003])
    for c in range(SO.N_CLASS):  # non-zero biases so that every term is exercised
        o = SO.HEAD + c * SO.CLASS_SIZE
        p[o + SO.O_B1:o + SO.O_W2] = rng.normal(0, 0.3, SO.N_HID)
        p[o + SO.O_B2:o + SO.O_W3] = rng.normal(0, 0.3, SO.N_HID)
        p[o + SO.O_TAIL] = 0.1
    return p


def test_cn_matches_d3_oracle(table, rcov):
    pos, Z, cell = small()
    o = d3_oracle.compute(pos, Z, cell, table, cutoff=30.0, cn_cutoff=30.0, atm_cutoff=0.0, damping=1, functional=0)
    cn, _ = SO.coordination(pos / SO.BOHR, Z, cell / SO.BOHR, rcov, 30.0)
    np.testing.assert_allclose(cn, o["cn"], rtol=1e-12, atol=1e-12)


def _c6_pair(table, Zi, Zj, cni, cnj):
    """C6(CN_i, CN_j) as the 25-term sum over the elements' valid references (CNref from the
    homonuclear entries), written out independently of the restatement."""
    c = np.asarray(table["c6ab"], np.float64)
    refs = lambda Zx: [(a, c[Zx, Zx, a, a, 1]) for a in range(table["nref"]) if c[Zx, Zx, a, a, 1] > -1]
    num = den = 0.0
    for a, ra in refs(Zi):
        for b, rb in refs(Zj):
            L = np.exp(-4.0 * (cni - ra) ** 2 - 4.0 * (cnj - rb) ** 2)
            num += c[Zi, Zj, a, b, 0] * L
            den += L
    return num / den


def test_isolated_pair_features(table):
    """Two atoms (H and O) far from every image: each sees one neighbour at r, weighted by
    sqrt(C6_HO) C6_jj^(1/4) at the atoms' CNs and the edge damping of r / R0_HO."""
    cell = np.eye(3) * 40.0
    r = 2.3
    pos = np.array([[1.0, 1.0, 1.0], [1.0 + r, 1.0, 1.0]])
    Z = np.array([1, 8])
    cn = np.array([0.37, 1.21])
    X, _ = SO.features(pos / SO.BOHR, Z, cell / SO.BOHR, cn, table)
    rb = r / SO.BOHR
    fc = 0.5 * (1 + np.cos(np.pi * rb / SO.RC_F))
    phi = np.exp(-(rb - SO.CENTRES) ** 2 / (2 * SO.SIGMA ** 2)) * fc
    c6 = _c6_pair(table, 1, 8, cn[0], cn[1])
    h = (1 + 6 * (4.0 / 3.0 * float(table["r0ab"][1, 8]) / rb) ** 16) ** -0.5
    u_HO = np.sqrt(c6) * _c6_pair(table, 8, 8, cn[1], cn[1]) ** 0.25 * h
    u_OH = np.sqrt(c6) * _c6_pair(table, 1, 1, cn[0], cn[0]) ** 0.25 * h
    np.testing.assert_allclose(X[0, 1 + 21:1 + 42], u_HO * phi, rtol=1e-12)  # H sees O (class 1)
    np.testing.assert_allclose(X[1, 1:22], u_OH * phi, rtol=1e-12)           # O sees H (class 0)
    assert np.count_nonzero(X[0, 1:]) == 21 and np.count_nonzero(X[1, 1:]) == 21


def test_neighbour_weight_derivatives_match_fd(table):
    """du_ij/dCN_i and du_ij/dCN_j of the restatement against central differences of u in CN."""
    Z = np.array([1, 8, 7, 29, 14])
    cn = np.array([0.6, 1.9, 2.4, 7.5, 3.3])
    i = np.array([0, 1, 2, 3, 4, 0, 3])
    j = np.array([1, 0, 3, 2, 1, 4, 3])
    u, du_i, du_j = SO.neighbour_weights(i, j, Z, cn, table)
    h = 1e-6
    for k in range(len(i)):
        for which, ref in ((i[k], du_i[k]), (j[k], du_j[k])):
            if which == i[k] and which == j[k]:
                continue  # a self pair moves both ends at once
            cp, cm = cn.copy(), cn.copy()
            cp[which] += h
            cm[which] -= h
            fd = (SO.neighbour_weights(i, j, Z, cp, table)[0][k] - SO.neighbour_weights(i, j, Z, cm, table)[0][k]) / (2 * h)
            assert abs(fd - ref) <= 1e-6 * max(1.0, abs(u[k])), (k, which)
    # and u against the written-out C6
    for k in range(len(i)):
        ref = np.sqrt(_c6_pair(table, Z[i[k]], Z[j[k]], cn[i[k]], cn[j[k]])) * \
            _c6_pair(table, Z[j[k]], Z[j[k]], cn[j[k]], cn[j[k]]) ** 0.25
        assert u[k] == pytest.approx(ref, rel=1e-12)


def test_input_gradient_matches_fd():
    p = random_params()
    rng = np.random.default_rng(4)
    X = rng.normal(0, 1, (9, SO.N_FEAT))
    cls = np.array([0, 1, 2, 0, 1, 2, 2, 1, 0])
    E, cache = SO.mlp_forward(p, X, cls)
    G = SO.mlp_input_grad(p, cache, len(X))
    h = 1e-6
    for a in range(len(X)):
        for f in (0, 1, 17, 40, 63):
            Xp, Xm = X.copy(), X.copy()
            Xp[a, f] += h
            Xm[a, f] -= h
            fd = (SO.mlp_forward(p, Xp, cls)[0][a] - SO.mlp_forward(p, Xm, cls)[0][a]) / (2 * h)
            assert abs(fd - G[a, f]) <= 1e-7 * max(1.0, abs(G[a, f]))


def test_param_gradient_matches_fd():
    p = random_params(seed=5)
    rng = np.random.default_rng(6)
    X = rng.normal(0, 1, (7, SO.N_FEAT))
    cls = np.array([0, 1, 2, 2, 1, 0, 2])
    y = rng.normal(0, 0.01, len(X))

    def loss(q):
        return 0.5 * np.sum((SO.mlp_forward(q, X, cls)[0] - y) ** 2)

    E, cache = SO.mlp_forward(p, X, cls)
    g = SO.mlp_param_grad(p, cache, E - y)
    h = 1e-6
    idx = []
    for c in range(SO.N_CLASS):
        o = SO.HEAD + c * SO.CLASS_SIZE
        idx += [o + SO.O_W1 + 5, o + SO.O_W1 + 64 * 128 - 3, o + SO.O_B1 + 7, o + SO.O_W2 + 300, o + SO.O_B2 + 1,
                o + SO.O_W3 + 9, o + SO.O_TAIL]
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (loss(pp) - loss(pm)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1e-3, abs(g[k])), k
    # normalisation and output-scale entries are not trained
    assert np.all(g[:SO.HEAD] == 0)
    for c in range(SO.N_CLASS):
        o = SO.HEAD + c * SO.CLASS_SIZE + SO.O_TAIL
        assert g[o + 1] == 0 and g[o + 2] == 0 and g[o + 3] == 0


def test_forces_and_stress_match_fd(table):
    # CNs inside the references' range (0.9-8.6), where the neighbour weights depend on them
    pos, Z, cell = small(seed=9, n=10)
    p = random_params(seed=2)
    o = SO.evaluate(p, pos, Z, cell, table, 25.0)
    h = 1e-5
    for a in (0, 3, 7):
        for k in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[a, k] += h
            pm[a, k] -= h
            fd = -(SO.evaluate(p, pp, Z, cell, table, 25.0)["energy"] - SO.evaluate(p, pm, Z, cell, table, 25.0)["energy"]) / (2 * h)
            assert abs(fd - o["forces"][a, k]) <= 1e-6 * max(1.0, np.max(np.abs(o["forces"])))
    V = abs(np.linalg.det(cell))
    # the D3 CN has a hard cutoff (a pair crossing it moves CN by ~1e-6): a strain step small
    # enough that no pair crosses it in this cell
    h = 1e-7
    for a, b in ((0, 0), (1, 2), (2, 1)):
        eps = np.zeros((3, 3))
        eps[a, b] = h
        ep = SO.evaluate(p, pos @ (np.eye(3) + eps), Z, cell @ (np.eye(3) + eps), table, 25.0)["energy"]
        em = SO.evaluate(p, pos @ (np.eye(3) - eps), Z, cell @ (np.eye(3) - eps), table, 25.0)["energy"]
        assert abs((ep - em) / (2 * h) / V - o["stress"][a, b]) <= 1e-6 * max(1e-3, np.max(np.abs(o["stress"])))
    # translation invariance: forces sum to zero
    assert np.max(np.abs(o["forces"].sum(0))) <= 1e-10 * max(1.0, np.max(np.abs(o["forces"])))


def test_bf16_rounding_and_emulated_network():
    """bf16(): nearest with ties to even; the bf16-rounded network stays within bf16 accuracy of
    the float64 one."""
    one = 1.0
    assert SO.bf16(one)[()] == 1.0
    assert SO.bf16(1 + 2 ** -8)[()] == 1.0                      # tie -> even (1.0)
    assert SO.bf16(1 + 3 * 2 ** -8)[()] == 1 + 2 ** -6          # tie -> even (upper)
    assert SO.bf16(1 + 2 ** -8 + 2 ** -12)[()] == 1 + 2 ** -7   # above the tie
    assert SO.bf16(-3.0)[()] == -3.0
    p = random_params(seed=11)
    rng = np.random.default_rng(12)
    X = rng.normal(0, 1, (50, SO.N_FEAT))
    cls = rng.integers(0, 3, 50)
    E64, cache = SO.mlp_forward(p, X, cls)
    G64 = SO.mlp_input_grad(p, cache, 50)
    E16, G16, _ = SO.mlp_bf16(p, X, cls)
    shift = np.array([SO.class_block(p, c)["shift"] for c in range(3)])[cls]
    assert np.max(np.abs(E16 - E64)) <= 3e-2 * np.max(np.abs(E64 - shift))
    assert np.max(np.abs(G16 - G64)) <= 5e-2 * np.max(np.abs(G64))
    # the emulated parameter gradient equals the float64 one to bf16 accuracy
    dl = rng.normal(0, 1, 50)
    _, _, g16 = SO.mlp_bf16(p, X, cls, dl=dl)
    g64 = SO.mlp_param_grad(p, cache, dl)
    assert np.max(np.abs(g16 - g64)) <= 5e-2 * np.max(np.abs(g64))


# ---------------------------------------------------------------------------- force matching

def test_tangent_matches_fd():
    """D = dE/dx . u equals the finite difference of the forward pass along u."""
    p = random_params(seed=21)
    rng = np.random.default_rng(22)
    X = rng.normal(0, 1, (9, SO.N_FEAT))
    U = rng.normal(0, 1, (9, SO.N_FEAT))
    cls = np.array([0, 1, 2, 0, 1, 2, 2, 1, 0])
    D, _ = SO.mlp_tangent(p, X, U, cls)
    h = 1e-6
    fd = (SO.mlp_forward(p, X + h * U, cls)[0] - SO.mlp_forward(p, X - h * U, cls)[0]) / (2 * h)
    assert np.max(np.abs(D - fd)) <= 1e-7 * max(1e-3, np.max(np.abs(D)))
    # and equals G . u from the input gradient
    E, cache = SO.mlp_forward(p, X, cls)
    G = SO.mlp_input_grad(p, cache, len(X))
    np.testing.assert_allclose(D, np.einsum("if,if->i", G, U), rtol=1e-12, atol=1e-15)


def test_tangent_param_gradient_matches_fd():
    """The reverse pass over the tangent: d(sum w D)/dparams equals finite differences."""
    p = random_params(seed=23)
    rng = np.random.default_rng(24)
    X = rng.normal(0, 1, (8, SO.N_FEAT))
    U = rng.normal(0, 1, (8, SO.N_FEAT))
    cls = np.array([0, 1, 2, 2, 1, 0, 2, 1])
    w = rng.normal(0, 1, len(X))

    def obj(q):
        return np.sum(w * SO.mlp_tangent(q, X, U, cls)[0])

    _, g = SO.mlp_tangent(p, X, U, cls, w)
    h = 1e-6
    idx = []
    for c in range(SO.N_CLASS):
        o = SO.HEAD + c * SO.CLASS_SIZE
        idx += [o + SO.O_W1 + 5, o + SO.O_W1 + 64 * 128 - 3, o + SO.O_W1 + 1000, o + SO.O_B1 + 7, o + SO.O_W2 + 300,
                o + SO.O_W2 + 128 * 128 - 1, o + SO.O_B2 + 1, o + SO.O_W3 + 9, o + SO.O_TAIL]
    scale = np.max(np.abs(g))
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (obj(pp) - obj(pm)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1e-3 * scale, abs(g[k])), k
    assert np.all(g[:SO.HEAD] == 0)
    for c in range(SO.N_CLASS):  # b3 does not enter a derivative; scale, shift are not trained
        o = SO.HEAD + c * SO.CLASS_SIZE + SO.O_TAIL
        assert g[o] == 0 and g[o + 1] == 0 and g[o + 2] == 0 and g[o + 3] == 0


def test_feature_tangent_matches_fd(table, rcov):
    """u = J dir equals the finite difference of the features (radial sums and CN) along dir."""
    pos, Z, cell = small(seed=9, n=10)
    rng = np.random.default_rng(25)
    dirv = rng.normal(0, 1, pos.shape)
    U = SO.feature_tangent(pos, Z, cell, table, 25.0, dirv)

    def feats(q):
        qb = q / SO.BOHR
        cn, _ = SO.coordination(qb, Z, cell / SO.BOHR, rcov, 25.0)
        return SO.features(qb, Z, cell / SO.BOHR, cn, table)[0]

    h = 1e-6
    fd = (feats(pos + h * dirv) - feats(pos - h * dirv)) / (2 * h)
    assert np.max(np.abs(U[:, 0] - fd[:, 0])) <= 1e-6 * max(1.0, np.max(np.abs(fd[:, 0])))
    assert np.max(np.abs(U[:, 1:] - fd[:, 1:])) <= 1e-6 * np.max(np.abs(fd[:, 1:]))
    # a rigid translation changes no feature
    Ut = SO.feature_tangent(pos, Z, cell, table, 25.0, np.tile([0.3, -1.0, 2.0], (len(Z), 1)))
    assert np.max(np.abs(Ut)) <= 1e-12 * np.max(np.abs(U))


def test_force_matching_gradient_matches_fd(table):
    """End to end: for L = sum |F(params) - F_ref|^2 of the stand-in on a small cell,
    dL/dparams = -2 sum_i d(G_i . u_i)/dparams with u = J (F - F_ref), which is mlp_tangent's
    gradient with w = -2 on the features' tangent along F - F_ref."""
    pos, Z, cell = small(seed=10, n=8)
    p = random_params(seed=26)
    rng = np.random.default_rng(27)
    o = SO.evaluate(p, pos, Z, cell, table, 25.0)
    F_ref = o["forces"] + rng.normal(0, np.std(o["forces"]), o["forces"].shape)

    def L(q):
        return np.sum((SO.evaluate(q, pos, Z, cell, table, 25.0)["forces"] - F_ref) ** 2)

    dF = o["forces"] - F_ref
    U = SO.feature_tangent(pos, Z, cell, table, 25.0, dF)
    cls = SO.element_class(Z)
    _, g = SO.mlp_tangent(p, o["features"], U, cls, np.full(len(Z), -2.0))
    # the identity behind it: sum_i G_i . u_i = -sum_i F_i . dF_i
    E, cache = SO.mlp_forward(p, o["features"], cls)
    G = SO.mlp_input_grad(p, cache, len(Z))
    assert abs(np.sum(G * U) + np.sum(o["forces"] * dF)) <= 1e-9 * np.sum(np.abs(o["forces"] * dF))
    present = sorted(set(cls.tolist()))
    idx = []
    for c in present:
        b = SO.HEAD + c * SO.CLASS_SIZE
        idx += [b + SO.O_W1 + 3, b + SO.O_W1 + 64 * 70 + 11, b + SO.O_B1 + 5, b + SO.O_W2 + 777, b + SO.O_B2 + 2,
                b + SO.O_W3 + 4]
    h = 1e-6
    scale = np.max(np.abs(g))
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (L(pp) - L(pm)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-5 * max(1e-3 * scale, abs(g[k])), k


def test_tangent_bf16_emulation_close_to_float64():
    p = random_params(seed=28)
    rng = np.random.default_rng(29)
    X = rng.normal(0, 1, (60, SO.N_FEAT))
    U = rng.normal(0, 1, (60, SO.N_FEAT))
    cls = rng.integers(0, 3, 60)
    w = rng.normal(0, 1, 60)
    D64, g64 = SO.mlp_tangent(p, X, U, cls, w)
    D16, g16 = SO.mlp_tangent_bf16(p, X, U, cls, w)
    assert np.max(np.abs(D16 - D64)) <= 5e-2 * np.max(np.abs(D64))
    assert np.max(np.abs(g16 - g64)) <= 5e-2 * np.max(np.abs(g64))
