"""CPU restatement of the learned three-body stand-in -- test infrastructure only.

Imported only by tests/ (and bench.py's accuracy check) as the checker for the HIP kernels in
lasp-d3_amd/csrc/d3_sur.hip; the product never loads it.  Plain float64 numpy, brute force over
periodic images, sized for small cells.

Model (no counterpart in the reference, which has only the exact ATM kernel,
src/d3_kernel.cu:1199-1509):
  features x_i (64): x_i[0] = CN_i (the D3 coordination number, src/d3_kernel.cu CN kernel with
    the handle's CN cutoff), x_i[1 + 21 c + m] = sum over neighbours j of element class c within
    RC_F of  u_ij exp(-(r_ij - mu_m)^2 / (2 sigma^2)) fc(r_ij),  sigma = 0.5 A, mu_m = 1 + 7 m / 20 A
    (m = 0..20), fc(r) = (1 + cos(pi r / RC_F)) / 2, RC_F = 9.5 A; classes: 0 = H, He;
    1 = Li..Ne; 2 = the rest.  Neighbour weight u_ij h_ij: u_ij = sqrt(C6_ij) C6_jj^(1/4) with the
    D3 C6 at the atoms' CNs (src/d3_kernel.cu:709-800 interpolation, in the engine's factorised
    form C6_ij = w_i^T C6ref w_j) -- for C6 close to the combination rule u_ij u_ik ~ sqrt(C6_ij
    C6_ik C6_jk), the three-body prefactor -- and h_ij = (1 + 6 ((4/3) R0_ij / r_ij)^16)^(-1/2) with
    R0 from r0ab, so that h_ij h_ik is the three-body damping (src/d3_kernel.cu:1386-1389) of an
    equilateral triangle of equal radii: an element enters through its C6 and its damping radii.
  per-atom network of the atom's own class c:  x' = (x - mean) * istd,
    h1 = silu(W1^T x' + b1) (64 -> 128),  h2 = silu(W2^T h1 + b2) (128 -> 128),
    E_i = scale_c (w3 . h2 + b3) + shift_c   [eV]
  energy E = sum_i E_i; forces and stress by the chain rule through the features
  (and through the CNs, on which x_i[0] and the weights u depend).
Force matching: with F = -J^T G (J: Jacobian of the features of every atom with respect to every
  position, G = dE_i/dx_i), the force loss sum |F - F_ref|^2 has the parameter gradient
  -2 sum_i d(G_i . u_i)/dparams at fixed u = J (F - F_ref) (feature_tangent); G_i . u_i is the
  network's derivative along u_i (mlp_tangent: forward tangent pass, then its reverse pass).
Parameter vector (float32 on the device, float64 here): mean[64], istd[64], then per class
  W1[64][128], b1[128], W2[128][128], b2[128], w3[128], (b3, scale, shift, pad).
The D3 CN here restates the oracle's C code (oracle/d3_oracle.cpp:156-170).
"""

from __future__ import annotations

import numpy as np

BOHR = 0.52917726  # Angstrom per bohr (src/d3.cu:50)
K1 = 16.0
N_CLASS, N_RBF, N_FEAT, N_HID = 3, 21, 64, 128
SIGMA = 0.5 / BOHR
CENTRES = (1.0 + 7.0 * np.arange(N_RBF) / (N_RBF - 1)) / BOHR
RC_F = 9.5 / BOHR

HEAD = 2 * N_FEAT
O_W1, O_B1 = 0, N_FEAT * N_HID
O_W2 = O_B1 + N_HID
O_B2 = O_W2 + N_HID * N_HID
O_W3 = O_B2 + N_HID
O_TAIL = O_W3 + N_HID
CLASS_SIZE = O_TAIL + 4
N_PARAMS = HEAD + N_CLASS * CLASS_SIZE


def element_class(Z) -> np.ndarray:
    Z = np.asarray(Z)
    return np.where(Z <= 2, 0, np.where(Z <= 10, 1, 2)).astype(np.int64)


def class_block(p: np.ndarray, c: int) -> dict:
    b = p[HEAD + c * CLASS_SIZE: HEAD + (c + 1) * CLASS_SIZE]
    return dict(W1=b[O_W1:O_B1].reshape(N_FEAT, N_HID), b1=b[O_B1:O_W2], W2=b[O_W2:O_B2].reshape(N_HID, N_HID),
                b2=b[O_B2:O_W3], w3=b[O_W3:O_TAIL], b3=b[O_TAIL], scale=b[O_TAIL + 1], shift=b[O_TAIL + 2])


def init_params(seed: int = 0, mean=None, istd=None, scale=None, shift=None) -> np.ndarray:
    """He-style random weights; normalisation and output scale from the arguments."""
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAMS)
    p[:N_FEAT] = 0.0 if mean is None else mean
    p[N_FEAT:HEAD] = 1.0 if istd is None else istd
    for c in range(N_CLASS):
        o = HEAD + c * CLASS_SIZE
        p[o + O_W1:o + O_B1] = rng.normal(0, np.sqrt(1.0 / N_FEAT), N_FEAT * N_HID)
        p[o + O_W2:o + O_B2] = rng.normal(0, np.sqrt(1.0 / N_HID), N_HID * N_HID)
        p[o + O_W3:o + O_TAIL] = rng.normal(0, np.sqrt(1.0 / N_HID), N_HID)
        p[o + O_TAIL + 1] = 1.0 if scale is None else scale[c]
        p[o + O_TAIL + 2] = 0.0 if shift is None else shift[c]
    return p


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def d2silu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 - s) * (2.0 + z * (1.0 - 2.0 * s))


# ---------------------------------------------------------------------------- geometry

def pair_images(pos_b: np.ndarray, cell_b: np.ndarray, rc: float):
    """All ordered pairs (i, j, image) with 0 < r <= rc: arrays i, j, d (= r_i - r_j - T, bohr), r."""
    n = len(pos_b)
    inv = np.linalg.inv(cell_b)
    heights = 1.0 / np.linalg.norm(inv, axis=0)
    reps = np.ceil(rc / heights).astype(int) + 1
    rng = [np.arange(-k, k + 1) for k in reps]
    T = np.stack(np.meshgrid(*rng, indexing="ij"), -1).reshape(-1, 3) @ cell_b
    I, J, D = [], [], []
    base = pos_b[:, None, :] - pos_b[None, :, :]
    for t in T:
        d = base - t
        r2 = np.einsum("ijk,ijk->ij", d, d)
        m = (r2 <= rc * rc) & (r2 > 1e-12)
        ii, jj = np.nonzero(m)
        I.append(ii)
        J.append(jj)
        D.append(d[ii, jj])
    i = np.concatenate(I)
    j = np.concatenate(J)
    d = np.concatenate(D)
    return i, j, d, np.sqrt(np.einsum("ij,ij->i", d, d))


def coordination(pos_b, Z, cell_b, rcov_of_Z, rc_cn):
    """D3 CN_i and the per-pair derivative dCN_i/dr_ij over the ordered pairs within rc_cn."""
    i, j, d, r = pair_images(pos_b, cell_b, rc_cn)
    rr = rcov_of_Z[Z[i]] + rcov_of_Z[Z[j]]
    e = np.exp(-K1 * (rr / r - 1.0))
    cn = np.bincount(i, weights=1.0 / (1.0 + e), minlength=len(Z))
    dc = -K1 * e * rr / ((1.0 + e) ** 2 * r * r)  # d/dr of 1/(1 + e)
    return cn, (i, j, d, r, dc)


def c6_weights(Z, cn, table):
    """Normalised Gaussian weights w[n, nref] of each atom's C6 references at its CN and dw/dCN
    (the engine's per-atom weights): w_a = exp(-4 (CN - CNref_a)^2) / sum_b exp(-4 (CN - CNref_b)^2)
    over the element's valid references (CNref from the homonuclear entry c6ab[Z][Z][a][a][1])."""
    c6ab = np.asarray(table["c6ab"], np.float64)
    nr = c6ab.shape[2]
    s = np.asarray(Z, np.int64) - table["zoff"]
    ref = c6ab[s, s][:, np.arange(nr), np.arange(nr), 1]
    valid = ref > -1
    dcn = np.asarray(cn, np.float64)[:, None] - ref
    ex = np.where(valid, -4.0 * dcn * dcn, -np.inf)
    e = np.where(valid, np.exp(ex - ex.max(1, keepdims=True)), 0.0)
    w = e / e.sum(1, keepdims=True)
    la = np.where(valid, -8.0 * dcn, 0.0)
    dw = w * (la - np.sum(w * la, 1, keepdims=True))
    return w, dw


def neighbour_weights(i, j, Z, cn, table):
    """u_ij = sqrt(C6_ij) q_j with q_j = C6_jj^(1/4), per ordered pair (i, j), and its derivatives
    with respect to CN_i and CN_j."""
    c6ab = np.asarray(table["c6ab"], np.float64)
    s = np.asarray(Z, np.int64) - table["zoff"]
    w, dw = c6_weights(Z, cn, table)
    ok = (c6ab[..., 1] > -1) & (c6ab[..., 2] > -1)
    ref = np.where(ok, c6ab[..., 0], 0.0)
    Cs = ref[s, s]  # (n, nr, nr)
    c6s = np.einsum("na,nab,nb->n", w, Cs, w)
    dc6s = 2.0 * np.einsum("na,nab,nb->n", dw, Cs, w)
    q = c6s ** 0.25
    dq = 0.25 * dc6s / c6s ** 0.75
    Cp = ref[s[i], s[j]]  # (npairs, nr, nr)
    c6 = np.einsum("pa,pab,pb->p", w[i], Cp, w[j])
    d_i = np.einsum("pa,pab,pb->p", dw[i], Cp, w[j])
    d_j = np.einsum("pa,pab,pb->p", w[i], Cp, dw[j])
    sc = np.sqrt(c6)
    u = sc * q[j]
    du_i = q[j] * d_i / (2.0 * sc)
    du_j = q[j] * d_j / (2.0 * sc) + sc * dq[j]
    return u, du_i, du_j


def edge_damping(r, R0):
    """h = (1 + 6 ((4/3) R0 / r)^16)^(-1/2) and dh/dr (0 for (4/3) R0 / r >= 100)."""
    t = np.minimum((4.0 / 3.0) * R0 / r, 100.0)
    s = t ** 16
    h = np.where(t < 100.0, (1.0 + 6.0 * s) ** -0.5, 0.0)
    return h, 48.0 * s * h ** 3 / r


def rbf(r):
    """phi_m(r) and dphi_m/dr, shape (len(r), 21)."""
    fc = 0.5 * (1.0 + np.cos(np.pi * r / RC_F))
    dfc = -0.5 * np.pi / RC_F * np.sin(np.pi * r / RC_F)
    u = r[:, None] - CENTRES[None, :]
    g = np.exp(-u * u / (2 * SIGMA * SIGMA))
    return g * fc[:, None], g * (dfc[:, None] - fc[:, None] * u / (SIGMA * SIGMA))


def features(pos_b, Z, cell_b, cn, table):
    """Features of every atom and, per ordered neighbour pair, what the derivatives need."""
    n = len(Z)
    i, j, d, r = pair_images(pos_b, cell_b, RC_F)
    phi, dphi = rbf(r)
    s = np.asarray(Z, np.int64) - table["zoff"]
    h, dh = edge_damping(r, np.asarray(table["r0ab"], np.float64)[s[i], s[j]])
    phi, dphi = phi * h[:, None], dphi * h[:, None] + phi * dh[:, None]  # the damped radial basis
    u, du_i, du_j = neighbour_weights(i, j, Z, cn, table)
    cls = element_class(Z)
    X = np.zeros((n, N_FEAT))
    X[:, 0] = cn
    col = 1 + N_RBF * cls[j][:, None] + np.arange(N_RBF)[None, :]
    np.add.at(X, (np.repeat(i, N_RBF), col.ravel()), (u[:, None] * phi).ravel())
    return X, dict(i=i, j=j, d=d, r=r, phi=phi, dphi=dphi, col=col, u=u, du_i=du_i, du_j=du_j)


# ---------------------------------------------------------------------------- network

def mlp_forward(p, X, cls):
    """Per-atom energies (eV) and the cache for the backward passes."""
    xs = (X - p[:N_FEAT]) * p[N_FEAT:HEAD]
    E = np.zeros(len(X))
    cache = []
    for c in range(N_CLASS):
        m = cls == c
        if not m.any():
            cache.append(None)
            continue
        w = class_block(p, c)
        z1 = xs[m] @ w["W1"] + w["b1"]
        h1 = silu(z1)
        z2 = h1 @ w["W2"] + w["b2"]
        h2 = silu(z2)
        y = h2 @ w["w3"] + w["b3"]
        E[m] = w["scale"] * y + w["shift"]
        cache.append(dict(m=m, xs=xs[m], z1=z1, h1=h1, z2=z2, h2=h2))
    return E, cache


def mlp_input_grad(p, cache, n):
    """dE_i/dx_i (n x 64)."""
    G = np.zeros((n, N_FEAT))
    for c, k in enumerate(cache):
        if k is None:
            continue
        w = class_block(p, c)
        dz2 = w["scale"] * w["w3"][None, :] * dsilu(k["z2"])
        dz1 = (dz2 @ w["W2"].T) * dsilu(k["z1"])
        G[k["m"]] = (dz1 @ w["W1"].T) * p[N_FEAT:HEAD]
    return G


def mlp_param_grad(p, cache, dLdE):
    """Gradient of L(E) with respect to the trainable parameters (W1, b1, W2, b2, w3, b3 of each
    class); dLdE: dL/dE_i per atom.  Normalisation, scale and shift entries get zero."""
    g = np.zeros_like(p)
    for c, k in enumerate(cache):
        if k is None:
            continue
        w = class_block(p, c)
        o = HEAD + c * CLASS_SIZE
        dy = w["scale"] * dLdE[k["m"]]
        g[o + O_W3:o + O_TAIL] = k["h2"].T @ dy
        g[o + O_TAIL] = dy.sum()
        dz2 = dy[:, None] * w["w3"][None, :] * dsilu(k["z2"])
        g[o + O_W2:o + O_B2] = (k["h1"].T @ dz2).ravel()
        g[o + O_B2:o + O_W3] = dz2.sum(0)
        dz1 = (dz2 @ w["W2"].T) * dsilu(k["z1"])
        g[o + O_W1:o + O_B1] = (k["xs"].T @ dz1).ravel()
        g[o + O_B1:o + O_W2] = dz1.sum(0)
    return g


def bf16(a) -> np.ndarray:
    """Rounds to bfloat16 (nearest, ties to even) through float32, returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def mlp_bf16(p, X, cls, dl=None):
    """The MFMA kernels' arithmetic (lasp-d3_amd/csrc/d3_mlp.hip): fp32 parameters; bf16 operands
    for x' = (x - mean) istd (formed in fp32), every weight matrix, h1, dz2 and dz1; biases, w3,
    SiLU values and accumulation wider.  Returns E, G = dE/dx and, when dl (dL/dE per row) is
    given, the gradient of L with respect to the parameters."""
    p32 = np.asarray(p, np.float32).astype(np.float64)
    X32 = np.asarray(X, np.float32)
    xs = bf16((X32 - p32[:N_FEAT].astype(np.float32)) * p32[N_FEAT:HEAD].astype(np.float32))
    n = len(X)
    E, G = np.zeros(n), np.zeros((n, N_FEAT))
    g = np.zeros_like(p32)
    for c in range(N_CLASS):
        m = cls == c
        if not m.any():
            continue
        w = class_block(p32, c)
        W1b, W2b = bf16(w["W1"]), bf16(w["W2"])
        z1 = xs[m] @ W1b + w["b1"]
        h1b = bf16(silu(z1))
        z2 = h1b @ W2b + w["b2"]
        h2 = silu(z2)
        E[m] = w["scale"] * (h2 @ w["w3"] + w["b3"]) + w["shift"]
        dy = w["scale"] * (np.ones(m.sum()) if dl is None else dl[m])
        dz2 = dy[:, None] * w["w3"][None, :] * dsilu(z2)
        dz2b = bf16(dz2)
        dz1 = (dz2b @ W2b.T) * dsilu(z1)
        dz1b = bf16(dz1)
        G[m] = (dz1b @ W1b.T) * p32[N_FEAT:HEAD]
        if dl is not None:
            o = HEAD + c * CLASS_SIZE
            g[o + O_W3:o + O_TAIL] = h2.T @ dy
            g[o + O_TAIL] = dy.sum()
            g[o + O_W2:o + O_B2] = (h1b.T @ dz2b).ravel()
            g[o + O_B2:o + O_W3] = dz2.sum(0)
            g[o + O_W1:o + O_B1] = (xs[m].T @ dz1b).ravel()
            g[o + O_B1:o + O_W2] = dz1.sum(0)
    return E, G, g


def mlp_tangent(p, X, U, cls, w=None):
    """D_i = dE_i/dx_i . u_i (the network's derivative along U, eV per unit of U) and, when w is
    given, the gradient of sum_i w_i D_i with respect to the trainable parameters (float64)."""
    istd = p[N_FEAT:HEAD]
    xs = (X - p[:N_FEAT]) * istd
    us = U * istd
    D = np.zeros(len(X))
    g = np.zeros_like(p)
    for c in range(N_CLASS):
        m = cls == c
        if not m.any():
            continue
        k = class_block(p, c)
        z1 = xs[m] @ k["W1"] + k["b1"]
        t1 = us[m] @ k["W1"]
        h1, s1 = silu(z1), dsilu(z1)
        h1t = s1 * t1
        z2 = h1 @ k["W2"] + k["b2"]
        t2 = h1t @ k["W2"]
        s2 = dsilu(z2)
        h2t = s2 * t2
        D[m] = k["scale"] * (h2t @ k["w3"])
        if w is None:
            continue
        o = HEAD + c * CLASS_SIZE
        gam = k["scale"] * w[m]
        g[o + O_W3:o + O_TAIL] = h2t.T @ gam
        dt2 = gam[:, None] * k["w3"][None, :] * s2
        dz2 = gam[:, None] * k["w3"][None, :] * d2silu(z2) * t2
        g[o + O_W2:o + O_B2] = (h1.T @ dz2 + h1t.T @ dt2).ravel()
        g[o + O_B2:o + O_W3] = dz2.sum(0)
        dh1, dh1t = dz2 @ k["W2"].T, dt2 @ k["W2"].T
        dt1 = dh1t * s1
        dz1 = dh1 * s1 + dh1t * d2silu(z1) * t1
        g[o + O_W1:o + O_B1] = (xs[m].T @ dz1 + us[m].T @ dt1).ravel()
        g[o + O_B1:o + O_W2] = dz1.sum(0)
    return D, g


def mlp_tangent_bf16(p, X, U, cls, w):
    """mlp_tangent with the MFMA kernel's roundings (lasp-d3_amd/csrc/d3_mlp.hip, tangent pass):
    bf16 operands for x', u' = u istd (formed in fp32), W1, W2, h1, h1' = silu'(z1) t1, dt2, dz2,
    dt1 and dz1; the rest wider.  Returns D and the parameter gradient of sum w D."""
    p32 = np.asarray(p, np.float32).astype(np.float64)
    X32, U32 = np.asarray(X, np.float32), np.asarray(U, np.float32)
    mu, istd = p32[:N_FEAT].astype(np.float32), p32[N_FEAT:HEAD].astype(np.float32)
    xs, us = bf16((X32 - mu) * istd), bf16(U32 * istd)
    D = np.zeros(len(X))
    g = np.zeros_like(p32)
    for c in range(N_CLASS):
        m = cls == c
        if not m.any():
            continue
        k = class_block(p32, c)
        W1b, W2b = bf16(k["W1"]), bf16(k["W2"])
        z1 = xs[m] @ W1b + k["b1"]
        t1 = us[m] @ W1b
        s1 = dsilu(z1)
        h1b, h1tb = bf16(silu(z1)), bf16(s1 * t1)
        z2 = h1b @ W2b + k["b2"]
        t2 = h1tb @ W2b
        s2 = dsilu(z2)
        h2t = s2 * t2
        D[m] = k["scale"] * (h2t @ k["w3"])
        o = HEAD + c * CLASS_SIZE
        gam = k["scale"] * w[m]
        g[o + O_W3:o + O_TAIL] = h2t.T @ gam
        dt2 = gam[:, None] * k["w3"][None, :] * s2
        dz2 = gam[:, None] * k["w3"][None, :] * d2silu(z2) * t2
        dt2b, dz2b = bf16(dt2), bf16(dz2)
        g[o + O_W2:o + O_B2] = (h1b.T @ dz2b + h1tb.T @ dt2b).ravel()
        g[o + O_B2:o + O_W3] = dz2.sum(0)
        dh1, dh1t = dz2b @ W2b.T, dt2b @ W2b.T
        dt1 = dh1t * s1
        dz1 = dh1 * s1 + dh1t * d2silu(z1) * t1
        g[o + O_W1:o + O_B1] = (xs[m].T @ bf16(dz1) + us[m].T @ bf16(dt1)).ravel()
        g[o + O_B1:o + O_W2] = dz1.sum(0)
    return D, g


# ---------------------------------------------------------------------------- energy and forces

def rcov_of(table):
    """rcov indexed by atomic number."""
    rc = np.asarray(table["rcov"], np.float64)
    return np.concatenate([np.zeros(table["zoff"]), rc]) if table["zoff"] else rc


def evaluate(p, pos_A, Z, cell_A, table, rc_cn, bf16_net=False):
    """Stand-in three-body energy (eV), per-atom energies, forces (eV/A), stress (eV/A^3, the
    strain derivative over the signed volume), features and dE/dCN (eV).  bf16_net: the network
    with the MFMA kernels' roundings (mlp_bf16) instead of float64."""
    Z = np.asarray(Z)
    pos_b = np.asarray(pos_A, float) / BOHR
    cell_b = np.asarray(cell_A, float) / BOHR
    n = len(Z)
    cls = element_class(Z)
    cn, (ci, cj, cd, cr, dc) = coordination(pos_b, Z, cell_b, rcov_of(table), rc_cn)
    X, a = features(pos_b, Z, cell_b, cn, table)
    i, j, d, r, col = a["i"], a["j"], a["d"], a["r"], a["col"]
    if bf16_net:
        E, G, _ = mlp_bf16(p, X, cls)
    else:
        E, cache = mlp_forward(p, X, cls)
        G = mlp_input_grad(p, cache, n)
    # per ordered pair (i, j, image) the features of i depend on r_ij and, through u_ij, on CN_i
    # and CN_j
    Gp = G[np.repeat(i, N_RBF), col.ravel()].reshape(-1, N_RBF)
    s = a["u"] * np.einsum("pm,pm->p", a["dphi"], Gp)
    A = np.einsum("pm,pm->p", a["phi"], Gp)
    dedcn = G[:, 0] + np.bincount(i, weights=A * a["du_i"], minlength=n) + \
        np.bincount(j, weights=A * a["du_j"], minlength=n)
    s_cn = dedcn[ci] * dc
    grad = np.zeros((n, 3))  # dE/dr (bohr)
    vir = np.zeros((3, 3))
    for (a, b, dd, rr, ss) in ((i, j, d, r, s), (ci, cj, cd, cr, s_cn)):
        v = (ss / rr)[:, None] * dd
        np.add.at(grad, a, v)
        np.add.at(grad, b, -v)
        vir += dd.T @ v
    V = np.linalg.det(cell_b)
    return dict(energy=E.sum(), atom_energy=E, forces=-grad / BOHR, stress=vir / V / BOHR ** 3, features=X,
                dedcn=dedcn, cn=cn)


def feature_tangent(pos_A, Z, cell_A, table, rc_cn, dirv):
    """u_i = sum_j (dx_i / dr_j) . dir_j: the change of every atom's features per unit step of
    the positions along dir (dir per A; u in feature units per A), through the radial sums, the
    neighbour weights (through CN_i and CN_j) and CN_i (x_i[0])."""
    Z = np.asarray(Z)
    pos_b = np.asarray(pos_A, float) / BOHR
    cell_b = np.asarray(cell_A, float) / BOHR
    dirv = np.asarray(dirv, float)
    n = len(Z)
    cn, (ci, cj, cd, cr, dc) = coordination(pos_b, Z, cell_b, rcov_of(table), rc_cn)
    _, a = features(pos_b, Z, cell_b, cn, table)
    i, j, d, r, col = a["i"], a["j"], a["d"], a["r"], a["col"]
    U = np.zeros((n, N_FEAT))
    # CN along dir (per A of step)
    pc = np.einsum("pk,pk->p", cd, dirv[ci] - dirv[cj]) / cr / BOHR
    dcn = np.bincount(ci, weights=dc * pc, minlength=n)
    U[:, 0] = dcn
    # dr_ij along dir (bohr per A of step): d_ij / r_ij . (dir_i - dir_j) / BOHR
    pr = np.einsum("pk,pk->p", d, dirv[i] - dirv[j]) / r / BOHR
    du = a["du_i"] * dcn[i] + a["du_j"] * dcn[j]
    v = a["u"][:, None] * a["dphi"] * pr[:, None] + a["phi"] * du[:, None]
    np.add.at(U, (np.repeat(i, N_RBF), col.ravel()), v.ravel())
    return U
