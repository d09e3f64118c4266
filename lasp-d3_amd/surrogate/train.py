"""Trains the per-atom network that stands in for the ATM three-body term, on the GPU, with the
library's own kernels (include/d3x_sur.h: d3x_sur_pack / d3x_sur_train_grad / d3x_sur_adam).

1. Training set from the exact kernel: seeded jittered frames of the 512-atom Si diamond cell
   (jitter N(0, 0.02 A), one seed per frame) and of the 4096-molecule water box (seeded boxes,
   each reused with per-frame jitter N(0, 0.02 A)), every frame with a seeded isotropic strain
   (default within +-3 %); the handle runs PBE-D3(BJ) with s6 = s8 = 0,
   so energies, forces and per-atom energies are those of the three-body term alone, in the mode
   that also returns the network's features.
2. The networks are fitted jointly to per-atom three-body energies and to the frames' three-body
   forces with Adam: each step takes random batches of atoms per class for the energy loss, and
   the force loss of one random training frame -- the network's forces on it from the engine,
   dF = F_net - F_exact, u = J dF through the features (d3x_sur_tangent) and the tangent-pass
   gradient of the network (d3x_sur_train_force_grad).  Forces normalised by each kind's RMS exact
   force.  --tie all (default): one network for every element class, whose first layer sees the
   three neighbour-class blocks of each radial centre through the same weights (gradients summed
   over the copies by d3x_sur_tie_grad), i.e. a function of CN and of the C6-weighted radial
   densities of all neighbours -- the training set holds H, O and Si centres only, and a block or
   a class network that no training atom exercises would answer arbitrarily for C, N or Cu;
   inputs normalised per radial centre by the statistics of the summed density, outputs by the
   label mean and spread of all atoms.  --tie classes: one network, neighbour blocks apart;
   --tie none: one network per element class (H; Li..Ne; heavier), each normalised by its own
   feature and label statistics.
3. Validation on held-out frames (every 20th): per-atom and per-frame energy errors, and through
   the engine: the judge's |dE| / atoms, force errors against the exact kernel, and the time of
   the three-body phase (exact kernel vs network).

Writes the parameter file (d3x_sur_load_file) and a JSON report.  PyTorch provides device memory
and the stream only.

usage: python lasp-d3_amd/surrogate/train.py [--si-frames 1000] [--water-frames 1000]
       [--water-bases 25] [--steps 20000] [--force-weight 1.0] [--out FILE] [--report FILE]
"""

from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "src"))

import d3_lib  # noqa: E402
import d3_structures as S  # noqa: E402
import d3_sur  # noqa: E402

RC2, RCCN, RC3 = float(np.sqrt(9000.0)), 40.0, 40.0
# {s6, s8_zero, s8_bj, sr6, sr8, a1, a2}: PBE-D3(BJ) with the two-body scale factors at zero
ATM_ONLY = np.array([0.0, 0.0, 0.0, 1.217, 1.0, 0.4289, 4.4407])
HOLDOUT = 20  # every 20th frame is held out


def log(*a):
    print("[train]", *a, flush=True)


def make_handle(Z, n, mode=d3_lib.D3X_ATM_EXACT_FEATURES, rc2=RC2):
    h = d3_lib.Handle(np.unique(Z), n, cutoff=rc2, cn_cutoff=RCCN, atm_cutoff=RC3, damping=1, functional=0,
                      precision=d3_lib.D3X_FP32)
    h.set_params(ATM_ONLY)
    h.set_atm_mode(mode)
    return h


def frame_stream(n_si, n_water, n_bases, strain):
    """Frames with a seeded isotropic strain in [-strain, strain] (positions and cell scaled
    together) so that the energy's dependence on density, hence the stress, is in the set."""
    rng = np.random.default_rng(7)
    for k in range(n_si):
        pos, Z, cell = S.si_diamond(4, jitter=0.02, seed=1000 + k)
        s = 1.0 + rng.uniform(-strain, strain)
        yield "si", k, pos * s, Z, cell * s
    bases = []
    for b in range(n_bases):
        bases.append(S.water_box(4096, seed=2000 + b))
        log(f"water base {b + 1}/{n_bases}")
    for k in range(n_water):
        pos, Z, cell = bases[k % n_bases]
        s = 1.0 + rng.uniform(-strain, strain)
        yield "water", n_si + k, (pos + rng.normal(0.0, 0.02, pos.shape)) * s, Z, cell * s


def build_set(args):
    n_atoms = args.si_frames * 512 + args.water_frames * 12288
    X = np.empty((n_atoms, d3_sur.NFEAT), np.float32)
    y = np.empty(n_atoms, np.float32)
    Zall = np.empty(n_atoms, np.int32)
    fid = np.empty(n_atoms, np.int32)
    frames = []
    handles = {}
    t0, r = time.time(), 0
    for kind, k, pos, Z, cell in frame_stream(args.si_frames, args.water_frames, args.water_bases, args.strain):
        if kind not in handles:
            handles[kind] = make_handle(Z, len(Z))
        h = handles[kind]
        h.set_atoms(pos, Z)
        h.set_cell(cell)
        e, f, s = h.compute()
        _, _, ea = h.atom_data()
        n = len(Z)
        X[r:r + n] = h.features()
        y[r:r + n] = ea
        Zall[r:r + n] = Z
        fid[r:r + n] = k
        frames.append(dict(kind=kind, index=k, start=r, n=n, energy=e, forces=f, pos=pos, Z=Z, cell=cell))
        r += n
        if len(frames) % 100 == 0:
            log(f"labelled {len(frames)} frames, {r} atoms, {time.time() - t0:.1f} s")
    for h in handles.values():
        h.close()
    return X, y, Zall, fid, frames, time.time() - t0


class ForceLoss:
    """Force-matching term of one frame per step.  The network's forces on the frame come from
    an engine handle in network mode (two-body scale factors at zero and a two-body cutoff below
    every bond, so the forces are the network's alone); u = J (F_net - F_exact) from
    d3x_sur_tangent; the gradient of weight * sum |F_net - F_exact|^2 / (3 n rms_kind^2) goes
    into grad through d3x_sur_train_force_grad, class by class."""

    def __init__(self, args, frames, Xd, cls, lib, st, work, grad, torch):
        self.torch, self.lib, self.st, self.work, self.grad = torch, lib, st, work, grad
        self.Xd, self.cls = Xd, cls
        self.weight = args.force_weight
        self.frames = [fr for fr in frames if fr["index"] % HOLDOUT != HOLDOUT - 1]
        self.rms = {}
        for kind in {fr["kind"] for fr in self.frames}:
            f = np.concatenate([fr["forces"].ravel() for fr in self.frames if fr["kind"] == kind])
            self.rms[kind] = float(np.sqrt(np.mean(f ** 2)))
        self.handles = {}
        self.rng = np.random.default_rng(args.seed + 1)
        dev = torch.device("cuda")
        nmax = max(fr["n"] for fr in self.frames)
        self.D = torch.zeros(nmax, device=dev)
        self.wbuf = torch.zeros(nmax, device=dev)
        self.loss_sum, self.count = 0.0, 0

    def handle(self, fr):
        h = self.handles.get(fr["kind"])
        if h is None:
            h = make_handle(fr["Z"], fr["n"], d3_lib.D3X_ATM_EXACT, rc2=1.0)
            self.handles[fr["kind"]] = h
        return h

    def step(self, P, pk):
        torch, lib, vp = self.torch, self.lib, (lambda t: C.c_void_p(t.data_ptr()))  # noqa: E731
        fr = self.frames[self.rng.integers(len(self.frames))]
        h = self.handle(fr)
        h.sur_load(P.cpu().numpy())
        h.set_atm_mode(d3_lib.D3X_ATM_SURROGATE, 0, 1e9)
        h.set_atoms(fr["pos"], fr["Z"])
        h.set_cell(fr["cell"])
        h.compute_device()  # not judged: the network alone
        _, F, _ = h.fetch()
        dF = F - fr["forces"]
        n = fr["n"]
        scale = self.weight / (3 * n * self.rms[fr["kind"]] ** 2)
        self.loss_sum += scale * float(np.sum(dF ** 2))
        self.count += 1
        U = torch.from_numpy(h.sur_tangent(dF)).to(P.device)
        a = fr["start"]
        cf = self.cls[a:a + n]
        chk = 0
        for c in np.unique(cf):
            idx = np.nonzero(cf == c)[0]
            ti = torch.from_numpy(idx).to(P.device)
            Xc = self.Xd[a:a + n].index_select(0, ti).contiguous()
            Uc = U.index_select(0, ti).contiguous()
            w = self.wbuf[:len(idx)]
            w.fill_(-2.0 * scale)
            chk |= lib.d3x_sur_train_force_grad(vp(P), vp(pk), int(c), len(idx), vp(Xc), vp(Uc), vp(w), vp(self.D),
                                                vp(self.work), vp(self.grad), self.st)
        return chk

    def take_loss(self):
        l = self.loss_sum / max(1, self.count)
        self.loss_sum, self.count = 0.0, 0
        return l

    def close(self):
        for h in self.handles.values():
            h.close()


def train(args, X, y, Zall, fid, frames):
    import torch
    lib = d3_lib.load()
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    vp = lambda t: C.c_void_p(t.data_ptr())  # noqa: E731
    cls = d3_sur.element_class(Zall)
    train_rows = (fid % HOLDOUT) != HOLDOUT - 1
    nr = d3_sur.NRBF
    Xt = X[train_rows].astype(np.float64)
    mean = Xt.mean(0)
    std = Xt.std(0)
    if args.tie == "all":
        # one weight per radial centre for the three neighbour blocks: normalise the summed density
        dens = Xt[:, 1:1 + nr] + Xt[:, 1 + nr:1 + 2 * nr] + Xt[:, 1 + 2 * nr:1 + 3 * nr]
        for b in range(3):
            mean[1 + b * nr:1 + (b + 1) * nr] = dens.mean(0) / 3.0
            std[1 + b * nr:1 + (b + 1) * nr] = dens.std(0)
    istd = 1.0 / np.maximum(std, 1e-3 * std.max())
    scale, shift = np.ones(3), np.zeros(3)
    present = []
    for c in range(3):
        m = train_rows & (cls == c)
        if m.any():
            present.append(c)
            shift[c] = y[m].mean()
            scale[c] = max(y[m].std(), 1e-7)
    if args.tie != "none":  # one network: one output scale
        shift[:] = y[train_rows].mean()
        scale[:] = max(y[train_rows].std(), 1e-7)
    p = d3_sur.init_params(args.seed, mean, istd, scale, shift)
    if args.tie != "none":  # identical copies (they stay identical: tied gradients, same Adam state)
        cs = d3_sur.CLASS_SIZE
        blk = p[d3_sur.HEAD:d3_sur.HEAD + cs].copy()
        if args.tie == "all":
            W1 = blk[:d3_sur.NFEAT * d3_sur.NHID].reshape(d3_sur.NFEAT, d3_sur.NHID)
            W1[1 + nr:1 + 2 * nr] = W1[1:1 + nr]
            W1[1 + 2 * nr:1 + 3 * nr] = W1[1:1 + nr]
        for c in range(3):
            p[d3_sur.HEAD + c * cs:d3_sur.HEAD + (c + 1) * cs] = blk
    dev = torch.device("cuda")
    P = torch.from_numpy(p).to(dev)
    Xd = torch.from_numpy(X).to(dev)
    yd = torch.from_numpy(y).to(dev)
    rows = {c: torch.from_numpy(np.nonzero(train_rows & (cls == c))[0]).to(dev) for c in present}
    B = args.batch
    grad, m1, m2 = torch.zeros_like(P), torch.zeros_like(P), torch.zeros_like(P)
    pk = torch.empty(lib.d3x_sur_packed_bytes(), dtype=torch.uint8, device=dev)
    nmax = max([B] + [fr["n"] for fr in frames])
    work = torch.empty(lib.d3x_sur_train_bytes(nmax), dtype=torch.uint8, device=dev)
    fl = None
    if args.force_weight > 0:
        fl = ForceLoss(args, frames, Xd, cls, lib, st, work, grad, torch)
        log("force rms per kind (eV/A):", json.dumps(fl.rms))
    E = torch.zeros(B, device=dev)
    loss = torch.zeros(1, device=dev)
    w = {c: torch.full((B,), 1.0 / (B * scale[c] ** 2), device=dev) for c in present}
    gen = torch.Generator(device=dev)
    gen.manual_seed(args.seed)
    curve = []
    t0 = time.time()
    for step in range(1, args.steps + 1):
        loss.zero_()
        chk = lib.d3x_sur_pack(vp(P), vp(pk), st)
        for c in present:
            sel = rows[c][torch.randint(0, len(rows[c]), (B,), device=dev, generator=gen)]
            Xb = Xd.index_select(0, sel).contiguous()
            yb = yd.index_select(0, sel).contiguous()
            chk |= lib.d3x_sur_train_grad(vp(P), vp(pk), c, B, vp(Xb), vp(yb), vp(w[c]), vp(E), vp(work), vp(grad),
                                          vp(loss), st)
        if fl is not None and step % args.force_every == 0:
            chk |= fl.step(P, pk)
        if args.tie != "none":
            chk |= lib.d3x_sur_tie_grad(vp(grad), 1 if args.tie == "all" else 0, st)
        frac = step / args.steps
        lr = args.lr * (0.5 * (1 + np.cos(np.pi * frac))) * min(1.0, step / 200) + 1e-6
        chk |= lib.d3x_sur_adam(vp(P), vp(grad), vp(m1), vp(m2), C.c_float(lr), C.c_float(0.9), C.c_float(0.999),
                                C.c_float(1e-8), step, st)
        if chk:
            raise RuntimeError(f"training kernels failed: {d3_lib.last_error()}")
        if step % 500 == 0 or step == 1:
            torch.cuda.synchronize()
            l = loss.item() / len(present)
            lf = fl.take_loss() if fl is not None else 0.0
            curve.append((step, l, lf))
            log(f"step {step} energy loss {l:.4e} force loss {lf:.4e} lr {lr:.2e} ({time.time() - t0:.1f} s)")
    torch.cuda.synchronize()
    if fl is not None:
        fl.close()
    p = P.cpu().numpy()
    # per-atom predictions for every row (held-out ones for validation)
    pred = np.empty(len(y), np.float32)
    chunk = 1 << 20
    Eall = torch.zeros(chunk, device=dev)
    cd = torch.from_numpy(cls).to(dev)
    lib.d3x_sur_pack(vp(P), vp(pk), st)
    for a in range(0, len(y), chunk):
        b = min(len(y), a + chunk)
        lib.d3x_sur_infer(vp(P), vp(pk), b - a, vp(Xd[a:b]), vp(cd[a:b]), vp(Eall), None, st)
        torch.cuda.synchronize()
        pred[a:b] = Eall[:b - a].cpu().numpy()
    return p, pred, curve, time.time() - t0, dict(mean=mean.tolist(), std=std.tolist(), scale=scale.tolist(),
                                                   shift=shift.tolist())


def engine_checks(p, cases):
    """Cells (kind, pos, Z, cell) through the engine: exact kernel vs the network (judge
    tolerance 1 meV/atom)."""
    out = []
    for kind, pos, Z, cell in cases:
        h = make_handle(Z, len(Z), d3_lib.D3X_ATM_EXACT)
        h.set_atoms(pos, Z)
        h.set_cell(cell)
        h.set_timing(True)
        e_x, f_x, s_x = h.compute()
        t_x = h.phase_times()[0]["atm"]
        ea_x = h.atom_data()[2]
        h.sur_load(p)
        h.set_atm_mode(d3_lib.D3X_ATM_SURROGATE, 0, 1e-3)
        h.set_timing(True)
        e_n, f_n, s_n = h.compute()  # judged: network and exact
        mode, fb, dev_ = h.sur_state()
        h.set_timing(True)
        h.set_atm_mode(d3_lib.D3X_ATM_SURROGATE, 0, 1e9)
        h.set_judge_force_tol(0.0)
        h.compute()  # judged with open tolerances: accepted
        h.set_timing(True)
        e_n2, f_n2, s_n2 = h.compute()  # not judged: the network alone
        t_n = h.phase_times()[0]["atm"]
        ea_n = h.atom_data()[2]
        h.close()
        n = len(Z)
        per_element = {}
        for z in np.unique(Z):
            m = Z == z
            d = 1e3 * (ea_n[m] - ea_x[m])
            per_element[int(z)] = dict(atoms=int(m.sum()), exact_mev_per_atom=float(1e3 * ea_x[m].mean()),
                                       mean_err_mev=float(d.mean()), rms_err_mev=float(np.sqrt(np.mean(d * d))),
                                       share_of_cell_err_mev_per_atom=float(d.sum() / n))
        out.append(dict(kind=kind, atoms=n, e_exact_mev_per_atom=1e3 * e_x / n,
                        e_net_mev_per_atom=1e3 * e_n2 / n, judge_dev_mev_per_atom=1e3 * dev_, fallback=fb,
                        mode_after=mode, force_rms_exact=float(np.sqrt(np.mean(f_x ** 2))),
                        force_rms_err=float(np.sqrt(np.mean((f_n2 - f_x) ** 2))),
                        stress_max_err=float(np.max(np.abs(s_n2 - s_x))), stress_max_exact=float(np.max(np.abs(s_x))),
                        atm_ms_exact=t_x, atm_ms_net=t_n, per_element=per_element))
        log("engine check", json.dumps(out[-1]))
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--si-frames", type=int, default=1000)
    ap.add_argument("--water-frames", type=int, default=1000)
    ap.add_argument("--water-bases", type=int, default=25)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--strain", type=float, default=0.03, help="isotropic strain range of the frames")
    ap.add_argument("--force-weight", type=float, default=1.0,
                    help="weight of the force loss (per kind normalised by the RMS exact force); 0: energies only")
    ap.add_argument("--force-every", type=int, default=1, help="force-loss frame every this many steps")
    ap.add_argument("--tie", choices=["all", "classes", "none"], default="all",
                    help="all: one network, neighbour-class blocks through shared weights; classes: one network; "
                         "none: one network per element class")
    ap.add_argument("--eval-random", type=int, default=0,
                    help="also check a random H/C/N/O/Si/Cu cell of this many atoms (0.08 atoms/A^3)")
    ap.add_argument("--out", default=os.path.join(HERE, "atm_surrogate.d3sur"))
    ap.add_argument("--report", default=None)
    args = ap.parse_args()
    # PyTorch has to bring up the device before the library does: in the other order its
    # runtime finds no GPU in this process
    import torch
    torch.zeros(1, device="cuda")
    X, y, Zall, fid, frames, t_lab = build_set(args)
    log(f"training set: {len(y)} atoms from {len(frames)} frames in {t_lab:.1f} s")
    p, pred, curve, t_train, stats = train(args, X, y, Zall, fid, frames)
    d3_sur.save(args.out, p)
    log(f"wrote {args.out}")
    held = (fid % HOLDOUT) == HOLDOUT - 1
    cls = d3_sur.element_class(Zall)
    per_class = {}
    for c in range(3):
        m = held & (cls == c)
        if m.any():
            per_class[c] = dict(atoms=int(m.sum()), rmse_mev=float(1e3 * np.sqrt(np.mean((pred[m] - y[m]) ** 2))),
                                label_std_mev=float(1e3 * y[m].std()), label_mean_mev=float(1e3 * y[m].mean()))
    per_frame = []
    for fr in frames:
        if fr["index"] % HOLDOUT != HOLDOUT - 1:
            continue
        a, n = fr["start"], fr["n"]
        per_frame.append((fr["kind"], 1e3 * abs(pred[a:a + n].sum() - y[a:a + n].sum()) / n))
    frame_err = {k: float(max(e for kk, e in per_frame if kk == k)) for k in {k for k, _ in per_frame}}
    cases, seen = [], set()
    for fr in frames:  # one held-out frame of each kind
        if fr["index"] % HOLDOUT == HOLDOUT - 1 and fr["kind"] not in seen:
            seen.add(fr["kind"])
            cases.append((fr["kind"], fr["pos"], fr["Z"], fr["cell"]))
    if args.eval_random:
        pos, Z, cell = S.random_cell(args.eval_random, seed=4242)
        cases.append((f"random{args.eval_random}", pos, Z, cell))
    checks = engine_checks(p, cases)
    report = dict(frames=len(frames), atoms=int(len(y)), label_seconds=t_lab, train_seconds=t_train,
                  steps=args.steps, batch=args.batch, force_weight=args.force_weight, force_every=args.force_every,
                  tie=args.tie,
                  curve=curve, heldout_per_class=per_class,
                  heldout_max_frame_error_mev_per_atom=frame_err, engine=checks, stats=stats)
    if args.report:
        with open(args.report, "w") as f:
            json.dump(report, f, indent=1)
    log(json.dumps(dict(per_class=per_class, frame_err=frame_err)))


if __name__ == "__main__":
    main()
