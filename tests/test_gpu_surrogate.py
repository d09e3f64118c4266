"""GPU checks of the learned three-body stand-in: the MFMA network kernels against the CPU
restatement with the kernels' bf16 roundings (oracle/surrogate_oracle.py mlp_bf16), Adam, a short
training run, the features and the full stand-in energy / forces / stress of the engine against
the oracle, and the judge / fallback behaviour."""

import ctypes as C

import numpy as np
import pytest

import d3_lib
import surrogate_oracle as SO
import systems

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev(a, dtype=torch.float32):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")


def vp(t):
    return C.c_void_p(t.data_ptr())


def stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def random_params(seed=1):
    rng = np.random.default_rng(seed)
    p = SO.init_params(seed, mean=rng.normal(0.5, 0.2, SO.N_FEAT), istd=rng.uniform(0.5, 2, SO.N_FEAT),
                       scale=[0.01, 0.02, 0.03], shift=[0.001, -0.002, 0.003])
    for c in range(SO.N_CLASS):
        o = SO.HEAD + c * SO.CLASS_SIZE
        p[o + SO.O_B1:o + SO.O_W2] = rng.normal(0, 0.3, SO.N_HID)
        p[o + SO.O_B2:o + SO.O_W3] = rng.normal(0, 0.3, SO.N_HID)
        p[o + SO.O_TAIL] = 0.1
    return p.astype(np.float32)


def packed(lib, P):
    pk = torch.empty(lib.d3x_sur_packed_bytes(), dtype=torch.uint8, device="cuda")
    assert lib.d3x_sur_pack(vp(P), vp(pk), stream()) == 0
    return pk


def test_infer_matches_bf16_oracle():
    lib = d3_lib.load()
    rng = np.random.default_rng(3)
    n = 1000
    X = rng.normal(0.5, 0.6, (n, SO.N_FEAT)).astype(np.float32)
    cls = rng.integers(0, 3, n).astype(np.int32)
    p = random_params()
    print("case sur-infer n=1000", flush=True)
    P = dev(p)
    pk = packed(lib, P)
    E = torch.zeros(n, device="cuda")
    G = torch.zeros((n, SO.N_FEAT), device="cuda")
    assert lib.d3x_sur_infer(vp(P), vp(pk), n, vp(dev(X)), vp(dev(cls, torch.int32)), vp(E), vp(G), stream()) == 0
    torch.cuda.synchronize()
    Eo, Go, _ = SO.mlp_bf16(p, X, cls)
    Eg, Gg = E.cpu().numpy(), G.cpu().numpy()
    # fp32 accumulation and a rare bf16 rounding flip of an activation: 1e-3 of the spread
    shift = np.array([SO.class_block(p.astype(np.float64), c)["shift"] for c in range(3)])[cls]
    assert np.max(np.abs(Eg - Eo)) <= 1e-3 * np.max(np.abs(Eo - shift))
    assert np.max(np.abs(Gg - Go)) <= 1e-2 * np.max(np.abs(Go))
    # and most entries much closer
    assert np.median(np.abs(Gg - Go)) <= 1e-4 * np.max(np.abs(Go))


def test_train_gradient_matches_bf16_oracle():
    lib = d3_lib.load()
    rng = np.random.default_rng(4)
    p = random_params(seed=7)
    P = dev(p)
    pk = packed(lib, P)
    grad = torch.zeros(d3_lib.SUR_NPARAMS, device="cuda")
    loss = torch.zeros(1, device="cuda")
    g_ref = np.zeros(d3_lib.SUR_NPARAMS)
    l_ref = 0.0
    for c, n in ((0, 333), (1, 1000), (2, 77)):
        print(f"case sur-train class={c} n={n}", flush=True)
        X = rng.normal(0.5, 0.6, (n, SO.N_FEAT)).astype(np.float32)
        y = rng.normal(0.0, 0.01, n).astype(np.float32)
        w = rng.uniform(0.5, 1.5, n).astype(np.float32) / n
        cls = np.full(n, c)
        E = torch.zeros(n, device="cuda")
        work = torch.empty(lib.d3x_sur_train_bytes(n), dtype=torch.uint8, device="cuda")
        assert lib.d3x_sur_train_grad(vp(P), vp(pk), c, n, vp(dev(X)), vp(dev(y)), vp(dev(w)), vp(E), vp(work),
                                      vp(grad), vp(loss), stream()) == 0
        Eo, _, _ = SO.mlp_bf16(p, X, cls)
        dl = w * (Eo - y)
        _, _, g = SO.mlp_bf16(p, X, cls, dl=dl)
        g_ref += g
        l_ref += np.sum(0.5 * w * (Eo - y) ** 2)
    torch.cuda.synchronize()
    gg = grad.cpu().numpy().astype(np.float64)
    assert abs(loss.item() - l_ref) <= 1e-4 * l_ref
    for c in range(3):
        o = SO.HEAD + c * SO.CLASS_SIZE
        for a, b in ((SO.O_W1, SO.O_B1), (SO.O_B1, SO.O_W2), (SO.O_W2, SO.O_B2), (SO.O_B2, SO.O_W3),
                     (SO.O_W3, SO.O_TAIL), (SO.O_TAIL, SO.O_TAIL + 1)):
            ref, got = g_ref[o + a:o + b], gg[o + a:o + b]
            assert np.max(np.abs(got - ref)) <= 2e-2 * np.max(np.abs(ref)) + 1e-12, (c, a)
            assert np.median(np.abs(got - ref)) <= 2e-3 * np.max(np.abs(ref)) + 1e-12, (c, a)
    assert np.all(gg[:SO.HEAD] == 0)


def mixed_cell(seed=21, n=48):
    s = systems.triclinic_mixed(n=n, seed=seed)
    return s["pos"], s["Z"], s["cell"]


RC = dict(cutoff=50.0, cn_cutoff=30.0, atm_cutoff=30.0, damping=1, functional=0)


def engine(Z, n, precision=d3_lib.D3X_FP64):
    return d3_lib.Handle(np.unique(Z), n, precision=precision, **RC)


@pytest.mark.parametrize("precision", [d3_lib.D3X_FP64, d3_lib.D3X_FP32])
def test_features_match_oracle(table, precision):
    pos, Z, cell = mixed_cell()
    print(f"case sur-features precision={precision}", flush=True)
    h = engine(Z, len(Z), precision)
    h.set_atm_mode(d3_lib.D3X_ATM_EXACT_FEATURES)
    h.set_atoms(pos, Z)
    h.set_cell(cell)
    h.compute()
    X = h.features().astype(np.float64)
    h.close()
    rcov = np.asarray(table["rcov"], np.float64)
    cn, _ = SO.coordination(pos / SO.BOHR, Z, cell / SO.BOHR, rcov, RC["cn_cutoff"])
    Xo, _ = SO.features(pos / SO.BOHR, Z, cell / SO.BOHR, cn, table)
    np.testing.assert_allclose(X[:, 0], Xo[:, 0], rtol=2e-6, atol=1e-6)
    np.testing.assert_allclose(X[:, 1:], Xo[:, 1:], rtol=1e-5, atol=1e-5 * np.max(Xo[:, 1:]))


def test_surrogate_energy_forces_stress_match_oracle(table):
    """The engine with the network in place of ATM = exact two-body (D3 oracle, no ATM) + the
    stand-in of the CPU restatement with the kernels' bf16 roundings."""
    import d3_oracle
    pos, Z, cell = mixed_cell(seed=22)
    p = random_params(seed=13)
    print("case sur-engine mixed48", flush=True)
    h = engine(Z, len(Z))
    h.sur_load(p)
    h.set_atm_mode(d3_lib.D3X_ATM_SURROGATE, 0, 1e9)  # the judge always accepts
    h.set_judge_force_tol(0.0)
    h.set_atoms(pos, Z)
    h.set_cell(cell)
    e, f, s = h.compute()
    mode, fb, dev_ = h.sur_state()
    h.close()
    assert mode == d3_lib.D3X_ATM_SURROGATE and fb == 0 and dev_ >= 0
    o2 = d3_oracle.compute(pos, Z, cell, table, cutoff=RC["cutoff"], cn_cutoff=RC["cn_cutoff"], atm_cutoff=0.0,
                           damping=1, functional=0)
    so = SO.evaluate(p.astype(np.float64), pos, Z, cell, table, RC["cn_cutoff"], bf16_net=True)
    assert abs((e - o2["energy"]) - so["energy"]) <= 1e-4 * np.sum(np.abs(so["atom_energy"])) + 1e-8
    df = f - o2["forces"]
    assert np.max(np.abs(df - so["forces"])) <= 2e-3 * np.max(np.abs(so["forces"])) + 1e-8
    ds = s - o2["stress"]
    assert np.max(np.abs(ds - so["stress"])) <= 2e-3 * np.max(np.abs(so["stress"])) + 1e-10


def test_judge_accepts_or_falls_back():
    pos, Z, cell = mixed_cell(seed=23)
    p = random_params(seed=14)
    print("case sur-judge", flush=True)
    h = engine(Z, len(Z))
    h.set_atoms(pos, Z)
    h.set_cell(cell)
    e_exact, f_exact, s_exact = h.compute()
    h.sur_load(p)
    # a network that cannot match to 1e-9 eV/atom: the judged first evaluation falls back
    h.set_atm_mode(d3_lib.D3X_ATM_SURROGATE, 0, 1e-9)
    e1, f1, s1 = h.compute()
    mode, fb, dev_ = h.sur_state()
    assert mode == d3_lib.D3X_ATM_EXACT and fb == 1 and dev_ > 1e-9
    assert abs(e1 - e_exact) <= 1e-12 * abs(e_exact)
    assert np.max(np.abs(f1 - f_exact)) <= 1e-10
    e2, _, _ = h.compute()  # stays on the exact kernel
    assert abs(e2 - e_exact) <= 1e-12 * abs(e_exact) and h.sur_state()[1] == 1
    # a generous energy tolerance, forces not judged: the network's results come back, judged
    # every second evaluation
    h.set_atm_mode(d3_lib.D3X_ATM_SURROGATE, 2, 1e9)
    h.set_judge_force_tol(0.0)
    e3, _, _ = h.compute()
    e4, _, _ = h.compute()
    assert h.sur_state()[0] == d3_lib.D3X_ATM_SURROGATE
    assert abs(e3 - e4) <= 1e-9 * abs(e3) and abs(e3 - e_exact) > 1e-9 * abs(e_exact)
    assert abs(h.sur_state()[2] - abs(e3 - e_exact) / len(Z)) <= 1e-6 * abs(e3 - e_exact) / len(Z) + 1e-12
    h.close()


def test_judge_force_criterion():
    """The judge's force criterion: the RMS of F_net - F_exact over the cell's force components is
    reported, and a force tolerance below it falls back to the exact kernel even when the energy
    tolerance accepts."""
    pos, Z, cell = mixed_cell(seed=24)
    p = random_params(seed=15)
    print("case sur-judge-forces", flush=True)
    h = engine(Z, len(Z))
    h.set_atoms(pos, Z)
    h.set_cell(cell)
    e_exact, f_exact, s_exact = h.compute()
    assert h.last_force_dev() < 0
    h.sur_load(p)
    h.set_atm_mode(d3_lib.D3X_ATM_SURROGATE, 0, 1e9)
    h.set_judge_force_tol(0.0)  # energies only: accepted
    e1, f1, _ = h.compute()
    assert h.sur_state()[:2] == (d3_lib.D3X_ATM_SURROGATE, 0)
    fdev = h.last_force_dev()
    assert fdev > 0
    assert fdev == pytest.approx(float(np.sqrt(np.mean((f1 - f_exact) ** 2))), rel=1e-6)
    # the same network, judged again with a force tolerance below its deviation
    h.set_atm_mode(d3_lib.D3X_ATM_SURROGATE, 0, 1e9)
    h.set_judge_force_tol(0.5 * fdev)
    e2, f2, s2 = h.compute()
    mode, fb, _ = h.sur_state()
    assert mode == d3_lib.D3X_ATM_EXACT and fb == 1
    assert abs(e2 - e_exact) <= 1e-12 * abs(e_exact)
    assert np.max(np.abs(f2 - f_exact)) <= 1e-10
    np.testing.assert_allclose(s2, s_exact, rtol=1e-10, atol=1e-14)
    # per-atom data of the handle are the exact kernel's after the fallback
    cn, g, ea = h.atom_data()
    assert abs(np.sum(ea) - e_exact) <= 1e-10 * abs(e_exact)
    # and a tolerance above it accepts
    h.set_atm_mode(d3_lib.D3X_ATM_SURROGATE, 0, 1e9)
    h.set_judge_force_tol(2.0 * fdev)
    e3, f3, _ = h.compute()
    assert h.sur_state()[0] == d3_lib.D3X_ATM_SURROGATE
    assert abs(e3 - e1) <= 1e-9 * abs(e1) and np.max(np.abs(f3 - f1)) <= 1e-9
    h.close()


def test_tie_grad_sums_the_copies():
    """d3x_sur_tie_grad against numpy: every class block's entries summed over the classes, and
    with blocks the radial rows of W1 summed over the three neighbour-class blocks too."""
    lib = d3_lib.load()
    rng = np.random.default_rng(31)
    g0 = rng.normal(0, 1, d3_lib.SUR_NPARAMS).astype(np.float32)
    for blocks in (0, 1):
        print(f"case sur-tie blocks={blocks}", flush=True)
        g = dev(g0)
        assert lib.d3x_sur_tie_grad(vp(g), blocks, stream()) == 0
        torch.cuda.synchronize()
        got = g.cpu().numpy().astype(np.float64)
        ref = g0.astype(np.float64).copy()
        cb = [ref[SO.HEAD + c * SO.CLASS_SIZE:SO.HEAD + (c + 1) * SO.CLASS_SIZE] for c in range(3)]
        tot = cb[0] + cb[1] + cb[2]
        W = tot[SO.O_W1:SO.O_B1].reshape(SO.N_FEAT, SO.N_HID)
        if blocks:
            s = W[1:22] + W[22:43] + W[43:64]
            W[1:22], W[22:43], W[43:64] = s, s, s
        for c in range(3):
            ref[SO.HEAD + c * SO.CLASS_SIZE:SO.HEAD + (c + 1) * SO.CLASS_SIZE] = tot
        np.testing.assert_allclose(got, ref, rtol=1e-6, atol=1e-6)


def test_adam_step_and_training_reduces_loss():
    lib = d3_lib.load()
    rng = np.random.default_rng(5)
    # one Adam step against numpy
    p = random_params(seed=9)
    g = rng.normal(0, 1e-3, d3_lib.SUR_NPARAMS).astype(np.float32)
    P, Gd = dev(p), dev(g)
    m, v = torch.zeros_like(P), torch.zeros_like(P)
    assert lib.d3x_sur_adam(vp(P), vp(Gd), vp(m), vp(v), C.c_float(1e-3), C.c_float(0.9), C.c_float(0.999),
                            C.c_float(1e-8), 1, stream()) == 0
    torch.cuda.synchronize()
    trainable = np.zeros(d3_lib.SUR_NPARAMS, bool)
    for c in range(3):
        o = SO.HEAD + c * SO.CLASS_SIZE
        trainable[o:o + SO.O_TAIL + 1] = True
    mm, vv = 0.1 * g, 0.001 * g.astype(np.float64) ** 2
    ref = p - np.where(trainable, 1e-3 * (mm / 0.1) / (np.sqrt(vv / 0.001) + 1e-8), 0)
    np.testing.assert_allclose(P.cpu().numpy(), ref, rtol=1e-5, atol=1e-7)
    assert np.all(Gd.cpu().numpy() == 0)
    # a short fit of a smooth target: the loss falls by 10x
    print("case sur-train-fit", flush=True)
    n = 4096
    X = rng.normal(0.0, 1.0, (n, SO.N_FEAT)).astype(np.float32)
    y = (0.01 * np.tanh(X[:, 0] + 0.5 * X[:, 5] - 0.3 * X[:, 40])).astype(np.float32)
    w = np.full(n, 1.0 / n, np.float32)
    p = SO.init_params(3, scale=[0.01, 0.01, 0.01]).astype(np.float32)
    P = dev(p)
    m, v = torch.zeros_like(P), torch.zeros_like(P)
    grad = torch.zeros_like(P)
    loss = torch.zeros(1, device="cuda")
    pk = torch.empty(lib.d3x_sur_packed_bytes(), dtype=torch.uint8, device="cuda")
    work = torch.empty(lib.d3x_sur_train_bytes(n), dtype=torch.uint8, device="cuda")
    E = torch.zeros(n, device="cuda")
    Xd, yd, wd = dev(X), dev(y), dev(w)
    losses = []
    for step in range(1, 301):
        loss.zero_()
        assert lib.d3x_sur_pack(vp(P), vp(pk), stream()) == 0
        assert lib.d3x_sur_train_grad(vp(P), vp(pk), 1, n, vp(Xd), vp(yd), vp(wd), vp(E), vp(work), vp(grad), vp(loss),
                                      stream()) == 0
        if step in (1, 300):
            torch.cuda.synchronize()
            losses.append(loss.item())
        assert lib.d3x_sur_adam(vp(P), vp(grad), vp(m), vp(v), C.c_float(3e-3), C.c_float(0.9), C.c_float(0.999),
                                C.c_float(1e-8), step, stream()) == 0
    assert losses[1] < 0.1 * losses[0], losses


# ---------------------------------------------------------------------------- force matching

def test_train_force_gradient_matches_bf16_oracle():
    """d3x_sur_train_force_grad: D = dE/dx . u per row and the parameter gradient of sum w D
    against the restatement with the kernel's bf16 roundings (mlp_tangent_bf16)."""
    lib = d3_lib.load()
    rng = np.random.default_rng(41)
    p = random_params(seed=17)
    P = dev(p)
    pk = packed(lib, P)
    grad = torch.zeros(d3_lib.SUR_NPARAMS, device="cuda")
    g_ref = np.zeros(d3_lib.SUR_NPARAMS)
    for c, n in ((0, 333), (1, 1000), (2, 77)):
        print(f"case sur-train-force class={c} n={n}", flush=True)
        X = rng.normal(0.5, 0.6, (n, SO.N_FEAT)).astype(np.float32)
        U = rng.normal(0.0, 0.3, (n, SO.N_FEAT)).astype(np.float32)
        w = rng.normal(0.0, 1.0, n).astype(np.float32) / n
        cls = np.full(n, c)
        D = torch.zeros(n, device="cuda")
        work = torch.empty(lib.d3x_sur_train_bytes(n), dtype=torch.uint8, device="cuda")
        assert lib.d3x_sur_train_force_grad(vp(P), vp(pk), c, n, vp(dev(X)), vp(dev(U)), vp(dev(w)), vp(D), vp(work),
                                            vp(grad), stream()) == 0
        torch.cuda.synchronize()
        Do, g = SO.mlp_tangent_bf16(p, X, U, cls, w)
        Dg = D.cpu().numpy()
        assert np.max(np.abs(Dg - Do)) <= 1e-2 * np.max(np.abs(Do))
        assert np.median(np.abs(Dg - Do)) <= 1e-4 * np.max(np.abs(Do))
        g_ref += g
    gg = grad.cpu().numpy().astype(np.float64)
    for c in range(3):
        o = SO.HEAD + c * SO.CLASS_SIZE
        for a, b in ((SO.O_W1, SO.O_B1), (SO.O_B1, SO.O_W2), (SO.O_W2, SO.O_B2), (SO.O_B2, SO.O_W3),
                     (SO.O_W3, SO.O_TAIL)):
            ref, got = g_ref[o + a:o + b], gg[o + a:o + b]
            assert np.max(np.abs(got - ref)) <= 2e-2 * np.max(np.abs(ref)) + 1e-12, (c, a)
            assert np.median(np.abs(got - ref)) <= 2e-3 * np.max(np.abs(ref)) + 1e-12, (c, a)
        assert gg[o + SO.O_TAIL] == 0  # b3 does not enter D
    assert np.all(gg[:SO.HEAD] == 0)


@pytest.mark.parametrize("precision", [d3_lib.D3X_FP64, d3_lib.D3X_FP32])
def test_feature_tangent_matches_oracle(table, precision):
    """d3x_sur_tangent (u = J dir through the radial sums and CN) against the restatement, whose
    own tangent equals finite differences of its features."""
    pos, Z, cell = mixed_cell(seed=24)
    rng = np.random.default_rng(42)
    dirv = rng.normal(0, 1, pos.shape)
    print(f"case sur-tangent precision={precision}", flush=True)
    h = engine(Z, len(Z), precision)
    h.set_atm_mode(d3_lib.D3X_ATM_EXACT_FEATURES)
    h.set_atoms(pos, Z)
    h.set_cell(cell)
    h.compute()
    U = h.sur_tangent(dirv).astype(np.float64)
    h.set_atoms(pos, Z)  # a geometry change invalidates the last evaluation's features
    with pytest.raises(d3_lib.D3Error):
        h.sur_tangent(dirv)
    h.close()
    Uo = SO.feature_tangent(pos, Z, cell, table, RC["cn_cutoff"], dirv)
    assert np.max(np.abs(U[:, 0] - Uo[:, 0])) <= 1e-5 * np.max(np.abs(Uo[:, 0]))
    assert np.max(np.abs(U[:, 1:] - Uo[:, 1:])) <= 1e-5 * np.max(np.abs(Uo[:, 1:]))


ATM_ONLY = np.array([0.0, 0.0, 0.0, 1.217, 1.0, 0.4289, 4.4407])  # PBE-D3(BJ), s6 = s8 = 0


@pytest.mark.parametrize("precision", [d3_lib.D3X_FP64, d3_lib.D3X_FP32])
def test_tangent_identity_full_water_box(precision):
    """At the full 12,288-atom water box: sum_i G_i . u_i = -sum_i F_i . dir_i, with u from
    d3x_sur_tangent, G from d3x_sur_infer on the engine's features and F the engine's network
    forces (two-body scale factors at zero, so that F is the network's alone)."""
    import d3_structures as S
    lib = d3_lib.load()
    pos, Z, cell = S.water_box(4096, seed=3001)
    n = len(Z)
    rng = np.random.default_rng(43)
    p = random_params(seed=19)
    print(f"case sur-tangent-identity water12288 precision={precision}", flush=True)
    h = d3_lib.Handle(np.unique(Z), n, cutoff=float(np.sqrt(9000.0)), cn_cutoff=40.0, atm_cutoff=40.0, damping=1,
                      functional=0, precision=precision)
    h.set_params(ATM_ONLY)
    h.sur_load(p)
    h.set_atm_mode(d3_lib.D3X_ATM_SURROGATE, 0, 1e9)
    h.set_atoms(pos, Z)
    h.set_cell(cell)
    h.compute_device()  # no judge: the network alone
    _, F, _ = h.fetch()
    X = h.features()
    dirv = F + rng.normal(0, np.std(F), F.shape)
    U = h.sur_tangent(dirv)
    h.close()
    cls = SO.element_class(Z).astype(np.int32)
    P = dev(p)
    pk = packed(lib, P)
    G = torch.zeros((n, SO.N_FEAT), device="cuda")
    assert lib.d3x_sur_infer(vp(P), vp(pk), n, vp(dev(X)), vp(dev(cls, torch.int32)), None, vp(G), stream()) == 0
    torch.cuda.synchronize()
    lhs = float(np.sum(G.cpu().numpy().astype(np.float64) * U))
    rhs = -float(np.sum(F * dirv))
    scale = float(np.sum(np.abs(F * dirv)))
    assert abs(lhs - rhs) <= 1e-4 * scale, (lhs, rhs, scale)


def test_force_matching_loop_reduces_force_loss():
    """The trainer's force step end to end on one Si 64 cell: network forces from the engine,
    u = J (F - F_exact) by d3x_sur_tangent, d3x_sur_train_force_grad with w = -2 / sum F_exact^2,
    Adam: 150 steps lower sum |F - F_exact|^2 at least 4x."""
    import d3_structures as S
    lib = d3_lib.load()
    pos, Z, cell = S.si_diamond(2, jitter=0.05, seed=5)
    n = len(Z)
    print(f"case sur-force-fit si{n}", flush=True)
    h = d3_lib.Handle(np.unique(Z), n, cutoff=1.0, cn_cutoff=40.0, atm_cutoff=40.0, damping=1, functional=0,
                      precision=d3_lib.D3X_FP32)
    h.set_params(ATM_ONLY)
    h.set_atm_mode(d3_lib.D3X_ATM_EXACT_FEATURES)
    h.set_atoms(pos, Z)
    h.set_cell(cell)
    _, F_ref, _ = h.compute()
    X = h.features()
    _, _, ea = h.atom_data()
    mean, std = X.mean(0), X.std(0)
    istd = 1.0 / np.maximum(std, 1e-3 * std.max())
    sc = max(float(np.std(ea)), 1e-6)
    p = SO.init_params(3, mean=mean, istd=istd, scale=[sc] * 3, shift=[float(ea.mean())] * 3).astype(np.float32)
    P = dev(p)
    m1, m2, grad = torch.zeros_like(P), torch.zeros_like(P), torch.zeros_like(P)
    pk = torch.empty(lib.d3x_sur_packed_bytes(), dtype=torch.uint8, device="cuda")
    work = torch.empty(lib.d3x_sur_train_bytes(n), dtype=torch.uint8, device="cuda")
    D = torch.zeros(n, device="cuda")
    Xd = dev(X)
    cls = int(SO.element_class(Z)[0])
    w = torch.full((n,), -2.0 / float(np.sum(F_ref ** 2)), device="cuda")
    losses = []
    for step in range(1, 151):
        h.sur_load(P.cpu().numpy())
        h.set_atm_mode(d3_lib.D3X_ATM_SURROGATE, 0, 1e9)
        h.compute_device()
        _, F, _ = h.fetch()
        dF = F - F_ref
        losses.append(float(np.sum(dF ** 2) / np.sum(F_ref ** 2)))
        U = dev(h.sur_tangent(dF))
        assert lib.d3x_sur_pack(vp(P), vp(pk), stream()) == 0
        assert lib.d3x_sur_train_force_grad(vp(P), vp(pk), cls, n, vp(Xd), vp(U), vp(w), vp(D), vp(work), vp(grad),
                                            stream()) == 0
        assert lib.d3x_sur_adam(vp(P), vp(grad), vp(m1), vp(m2), C.c_float(2e-3), C.c_float(0.9), C.c_float(0.999),
                                C.c_float(1e-8), step, stream()) == 0
        torch.cuda.synchronize()
    h.close()
    print(f"  relative force loss {losses[0]:.3e} -> {losses[-1]:.3e}", flush=True)
    assert losses[-1] < 0.25 * losses[0], (losses[0], losses[-1])
