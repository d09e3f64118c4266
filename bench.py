#!/usr/bin/env python3
"""Throughput benchmark of the D3 engine on one MI355X.

A step is one full evaluation -- energy, forces and 3x3 stress with two-body, CN, ATM
three-body and the dE/dCN chain rule -- of one synthetic periodic structure whose coordinates
are already resident in HBM (d3x_compute_device on a handle set up beforehand).  The headline
value is the FP32 (compensated) path in atoms/s; the plain FP64 path is timed on the same input
for the "vs plain FP64" comparison, and the FP32 error against it is reported.

Workloads (seeded generators in lasp-d3_amd/src/d3_structures.py; PBE-D3(BJ), two-body cutoff
sqrt(9000) bohr, CN and ATM cutoff 40 bohr):
  cu1m   fcc Cu 63^3, 1,000,188 atoms, a = 3.615 A, jitter N(0, 0.05 A) (default: the cell the
         FP32-vs-FP64 target is quoted on)
  water  4096 rigid H2O, 12,288 atoms, 49.7 A cubic cell
  si512  diamond Si 4x4x4, 512 atoms
  batch  256 triclinic cells of 50-5000 atoms (~300k atoms), one batched evaluation

Prints ONE JSON line.  roofline: the dominant kernel's algorithmic FP32 work per launch
(work model: 17 flops per unique CN pair, 193 per unique two-body pair, 330 per unique
triangle, 45 per unique CN pair for the chain rule) over its HIP-event-timed average launch
duration, against the 157.3 TFLOP/s FP32 vector peak.  cpu_baseline: the FP64 OpenMP oracle
timed on a bounded random sample of centre atoms on this host (scaled per atom).
"""

from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _sub in ("lasp-d3_amd/src", "oracle"):
    sys.path.insert(0, os.path.join(ROOT, _sub))

import numpy as np  # noqa: E402

import d3_lib  # noqa: E402
import d3_structures as S  # noqa: E402

RC2 = float(np.sqrt(9000.0))
RCCN = 40.0
RC3 = 40.0
PEAK_FP32 = 157.3  # TFLOP/s, MI355X vector FP32
FLOPS = dict(cn=17.0, two_body=193.0, atm=330.0, chain=45.0)
METRIC = "atoms/s E+F+stress (2b+3b) vs plain FP64 GPU; fraction of FP32 VALU peak"


class HipEvents:
    """hipEvent timing on an arbitrary stream through the HIP runtime's C API."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so")

    def create(self):
        ev = C.c_void_p()
        assert self.hip.hipEventCreate(C.byref(ev)) == 0
        return ev

    def record(self, ev, stream):
        assert self.hip.hipEventRecord(ev, C.c_void_p(stream)) == 0

    def elapsed_ms(self, a, b):
        assert self.hip.hipEventSynchronize(b) == 0
        ms = C.c_float()
        assert self.hip.hipEventElapsedTime(C.byref(ms), a, b) == 0
        return float(ms.value)


def workload(name: str):
    if name == "water":
        pos, Z, cell = S.water_box(4096, seed=0)
        return dict(kind="single", pos=pos, Z=Z, cell=cell, desc="water-4096H2O")
    if name == "si512":
        pos, Z, cell = S.si_diamond(4, jitter=0.02, seed=0)
        return dict(kind="single", pos=pos, Z=Z, cell=cell, desc="si-diamond-512")
    if name == "cu1m":
        pos, Z, cell = S.fcc_cu(63, jitter=0.05, seed=0)
        return dict(kind="single", pos=pos, Z=Z, cell=cell, desc="fcc-cu-1000188")
    if name == "batch":
        counts, pos, Z, cells = S.triclinic_batch(256, seed=0)
        return dict(kind="batch", counts=counts, pos=pos, Z=Z, cells=cells, desc="triclinic-batch-256")
    raise SystemExit(f"unknown workload {name}")


def make_handle(w, precision, graph=False, autotune=False):
    n = len(w["Z"])
    h = d3_lib.Handle(np.unique(w["Z"]), n, cutoff=RC2, cn_cutoff=RCCN, atm_cutoff=RC3, damping=1, functional=0,
                      precision=precision)
    if w["kind"] == "single":
        h.set_atoms(w["pos"], w["Z"])
        h.set_cell(w["cell"])
        h.compute()  # sizes scratch (retries a neighbour-list overflow) and leaves inputs resident
    else:
        h.compute_batch(w["counts"], w["pos"], w["Z"], w["cells"])
    if autotune:
        os.environ.setdefault("D3_TUNE_LOG", "1")  # a line per timed candidate (keeps long sweeps visible)
        h.autotune()  # times the candidates on this input and keeps the fastest
    if graph:
        h.set_graph(True)
        h.compute_device()  # captures; the timed steps replay
    return h


def fetch(h, w):
    if w["kind"] == "single":
        return h.fetch()
    e = np.zeros(len(w["counts"]))
    f = np.zeros((len(w["Z"]), 3))
    s = np.zeros((len(w["counts"]), 9))
    rc = h.lib.d3x_fetch_results(h.h, d3_lib.ptr(e, C.c_double), d3_lib.ptr(f, C.c_double),
                                 d3_lib.ptr(s, C.c_double))
    if rc:
        raise RuntimeError(d3_lib.last_error())
    return e, f, s


def time_steps(h, ev, steps, warmup):
    for _ in range(warmup):
        h.compute_device()
    st = h.stream()
    a, b = ev.create(), ev.create()
    ev.hip.hipStreamSynchronize(C.c_void_p(st))
    t0 = time.perf_counter()
    ev.record(a, st)
    for _ in range(steps):
        h.compute_device()
    ev.record(b, st)
    ms = ev.elapsed_ms(a, b)
    wall = time.perf_counter() - t0
    return ms / steps, wall


def host_to_host(w, runs=5):
    """The timed region of a caller with host buffers: positions, types and cell in host memory ->
    energy, forces and stress in host memory (d3x_set_atoms_f64, d3x_set_cell_f64, d3x_compute_f64 on
    an FP32 handle), wall clock, one warm-up then the median of `runs`."""
    n = len(w["Z"])
    h = d3_lib.Handle(np.unique(w["Z"]), n, cutoff=RC2, cn_cutoff=RCCN, atm_cutoff=RC3, damping=1, functional=0,
                      precision=d3_lib.D3X_FP32)
    times = []
    for k in range(runs + 1):
        t0 = time.perf_counter()
        h.set_atoms(w["pos"], w["Z"])
        h.set_cell(w["cell"])
        h.compute()
        if k:
            times.append(time.perf_counter() - t0)
    h.close()
    ms = 1e3 * float(np.median(times))
    return dict(ms_per_step=ms, atoms_per_s=n / (ms * 1e-3), runs=runs,
                note="host positions in, host E/F/stress out, PCIe and host preparation included")


def phase_profile(h, runs=2):
    h.set_timing(True)
    for _ in range(runs):
        h.compute_device()
    ph, n = h.phase_times()
    h.set_timing(False)
    return {k: v / max(n, 1) for k, v in ph.items()}


def hbm_used_gb(ev):
    """Device memory in use by this process and the runtime, from hipMemGetInfo (GB)."""
    free, total = C.c_size_t(), C.c_size_t()
    if ev.hip.hipMemGetInfo(C.byref(free), C.byref(total)) != 0:
        return None
    return (total.value - free.value) / 1e9


SUR_WEIGHTS = os.path.join(ROOT, "lasp-d3_amd", "surrogate", "atm_surrogate.d3sur")


def profiled_traffic(kernel: str, workload_desc: str):
    """HBM bytes per launch of `kernel` on the workload, from the counter summary committed under
    profiles/ (rocprofv3 FETCH_SIZE and WRITE_SIZE runs scaled by a plain copy of known size):
    the line 'traffic_bytes_per_launch: X' of profiles/<kernel>_<workload>_traffic.md, or None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload_desc}_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def surrogate_leg():
    """The three-body network against the exact ATM kernel (FP32 handles, same cutoffs), on the
    water box (its training distribution; the box itself is not in the training set) and on the
    100,000-atom random H/C/N/O/Si/Cu cell at 0.08 atoms/A^3 (C, N and Cu are in no training
    frame): three-body phase time of each (HIP events, average of 3 evaluations), whole-step
    time, the judge's |dE| / atoms against the exact kernel with the 1 meV/atom energy criterion
    and whether it fell back, the RMS force deviation (full D3 both: the difference is the
    three-body term's) and whether the judge's default force criterion (5e-3 eV/A RMS) also
    accepts, and per element the mean and RMS error of the per-atom energies."""
    cases = []
    for name, (pos, Z, cell) in (("water-4096H2O", S.water_box(4096, seed=0)),
                                  ("random-100000", S.random_cell(100000, seed=4242))):
        n = len(Z)
        print(f"[bench] surrogate leg {name}", file=sys.stderr, flush=True)
        h = d3_lib.Handle(np.unique(Z), n, cutoff=RC2, cn_cutoff=RCCN, atm_cutoff=RC3, damping=1, functional=0,
                          precision=d3_lib.D3X_FP32)
        h.set_atoms(pos, Z)
        h.set_cell(cell)
        e_x, f_x, _ = h.compute()
        ea_x = h.atom_data()[2]
        px = phase_profile(h, runs=3)
        h.sur_load_file(SUR_WEIGHTS)
        # the energy criterion of the target: |dE| / atoms <= 1 meV
        h.set_atm_mode(d3_lib.D3X_ATM_SURROGATE, 0, 1e-3)
        h.set_judge_force_tol(0.0)
        e_n, f_n, _ = h.compute()  # judged against the exact kernel (its results after a fallback)
        mode, fb, dev = h.sur_state()
        fdev = h.last_force_dev()
        # the same judgement with the default force criterion as well
        h.set_atm_mode(d3_lib.D3X_ATM_SURROGATE, 0, 1e-3)
        h.set_judge_force_tol(5e-3)
        h.compute()
        fb_force = h.sur_state()[0] != d3_lib.D3X_ATM_SURROGATE
        # the network alone (judge accepting), for its time and per-atom energies
        h.set_atm_mode(d3_lib.D3X_ATM_SURROGATE, 0, 1e9)
        h.set_judge_force_tol(0.0)
        h.compute()
        pn = phase_profile(h, runs=3)
        ea_n = h.atom_data()[2]
        h.close()
        per_element = {}
        for z in np.unique(Z):
            m = Z == z
            d = 1e3 * (ea_n[m] - ea_x[m])
            per_element[int(z)] = dict(atoms=int(m.sum()), exact_mev_per_atom=float(1e3 * ea_x[m].mean()),
                                       mean_err_mev=float(d.mean()), rms_err_mev=float(np.sqrt(np.mean(d * d))),
                                       share_of_cell_err_mev_per_atom=float(d.sum() / n))
        cases.append(dict(workload=name, atoms=n, atm_ms_exact=px["atm"], atm_ms_network=pn["atm"],
                          atm_speedup=px["atm"] / max(pn["atm"], 1e-9), step_ms_exact=sum(px.values()),
                          step_ms_network=sum(pn.values()), judge_dev_mev_per_atom=1e3 * dev,
                          fell_back=bool(fb), mode_after=mode, e_exact_ev=e_x, e_returned_ev=e_n,
                          force_rms_exact_ev_per_a=float(np.sqrt(np.mean(f_x ** 2))),
                          force_rms_dev_ev_per_a=fdev,
                          force_rms_err_ev_per_a=float(np.sqrt(np.mean((f_n - f_x) ** 2))),
                          fell_back_with_force_check=bool(fb_force), per_element=per_element))
    return dict(weights=os.path.relpath(SUR_WEIGHTS, ROOT), tolerance_mev_per_atom=1.0,
                force_check_ev_per_a=5e-3, cases=cases)


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(w, table, cn, g, target_s=15.0, target_1t_s=5.0):
    """FP64 oracle on a bounded random sample of centre atoms: all host threads of this job, then
    one thread on a smaller sample."""
    import d3_oracle
    if w["kind"] != "single":
        return None
    n = len(w["Z"])
    # every CPU this job may run on (the box sets OMP_NUM_THREADS to its CPU share), threads kept
    # close to their cores
    os.environ.setdefault("OMP_PROC_BIND", "close")
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0))
    rng = np.random.Generator(np.random.PCG64(1234))
    kw = dict(cutoff=RC2, cn_cutoff=RCCN, atm_cutoff=RC3, damping=1, functional=0)
    probe = rng.choice(n, size=min(n, threads), replace=False)
    t1, _ = d3_oracle.time_centers(w["pos"], w["Z"], w["cell"], table, probe, cn, g, nthreads=threads, **kw)
    reps = max(1, int(target_s / max(t1, 1e-6)))
    m = int(min(n, threads * reps))
    sample = rng.choice(n, size=m, replace=False)
    t, _ = d3_oracle.time_centers(w["pos"], w["Z"], w["cell"], table, sample, cn, g, nthreads=threads, **kw)
    # one thread: as many centres as take about target_1t_s at the per-centre cost just measured
    m1 = int(max(1, min(n, target_1t_s * m / (t * threads))))
    s1 = rng.choice(n, size=m1, replace=False)
    ts, _ = d3_oracle.time_centers(w["pos"], w["Z"], w["cell"], table, s1, cn, g, nthreads=1, **kw)
    print(f"[bench] cpu baseline: {threads} threads {m / t:.1f} atoms/s, 1 thread {m1 / ts:.2f} atoms/s",
          file=sys.stderr, flush=True)
    return dict(value=m / t, unit="atoms/s", cores=threads, kind="port",
                sample=f"{m} random centre atoms of {n} (all four passes per centre, FP64 oracle, "
                       f"{t:.1f} s); throughput scales per atom because every atom's work is one centre's",
                value_1_thread=m1 / ts, sample_1_thread=f"{m1} random centre atoms, {ts:.1f} s",
                host_cpus=os.cpu_count(), cpu_model=cpu_model())


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="cu1m", choices=["water", "si512", "cu1m", "batch"])
    ap.add_argument("--fp64-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-surrogate", action="store_true", help="skip the three-body network leg")
    ap.add_argument("--graph", action="store_true", help="replay the evaluation as a captured hipGraph")
    ap.add_argument("--autotune", action="store_true", help="autotune bin width and three-body grid first")
    ap.add_argument("--only-surrogate", action="store_true",
                    help="run only the three-body network leg and print its JSON (for counter runs)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only copy BYTES device-to-device three times with a plain copy kernel "
                         "(reference for scaling FETCH_SIZE / WRITE_SIZE counter readings)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        import torch  # a plain device copy kernel of known size
        x = torch.ones(args.copy_calibration // 4, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": args.copy_calibration, "copies": 3}))
        return
    if args.only_surrogate:
        print(json.dumps({"surrogate": surrogate_leg()}))
        return

    w = workload(args.workload)
    n = len(w["Z"])
    ev = HipEvents()
    print(f"[bench] workload {w['desc']} atoms={n}", file=sys.stderr, flush=True)

    hbm0 = hbm_used_gb(ev)
    h32 = make_handle(w, d3_lib.D3X_FP32, graph=args.graph, autotune=args.autotune)
    tuning32 = h32.tuning()
    variant32 = h32.atm_variant() if RC3 > 0 else None
    hbm32 = hbm_used_gb(ev)
    ms32, wall32 = time_steps(h32, ev, args.steps, args.warmup)
    e32, f32, s32 = fetch(h32, w)
    print(f"[bench] fp32 {ms32:.2f} ms/step", file=sys.stderr, flush=True)
    prof32 = phase_profile(h32)
    counts = h32.counts(len(w["counts"]) if w["kind"] == "batch" else 1).sum(axis=0)

    h64 = make_handle(w, d3_lib.D3X_FP64, graph=args.graph, autotune=args.autotune)
    k64 = args.fp64_steps or args.steps
    ms64, _ = time_steps(h64, ev, k64, 1)
    e64, f64, s64 = fetch(h64, w)
    print(f"[bench] fp64 {ms64:.2f} ms/step", file=sys.stderr, flush=True)
    prof64 = phase_profile(h64, runs=1)
    if w["kind"] == "single":
        cn64, g64, _ = h64.atom_data()

    e32a, e64a = np.atleast_1d(e32), np.atleast_1d(e64)
    err = dict(energy_rel=float(np.max(np.abs(e32a - e64a) / np.abs(e64a))),
               force_max_abs=float(np.max(np.abs(f32 - f64))),
               force_rel_to_max=float(np.max(np.abs(f32 - f64)) / max(np.max(np.abs(f64)), 1e-300)),
               stress_max_abs=float(np.max(np.abs(np.asarray(s32) - np.asarray(s64)))))

    # work model: unique pairs / triangles from the GPU's own counts
    uniq = dict(two_body=counts[0] / 2, cn=counts[1] / 2, chain=counts[1] / 2, atm=counts[2])
    flops = {k: FLOPS[k] * uniq[k] for k in FLOPS}
    dom = max(("cn", "two_body", "atm", "chain"), key=lambda k: prof32[k])
    achieved = flops[dom] / (prof32[dom] * 1e-3) / 1e12
    total_flops = sum(flops.values())

    hbm = dict(fp32_handle_gb=None if hbm0 is None or hbm32 is None else hbm32 - hbm0,
               both_handles_gb=None if hbm0 is None else hbm_used_gb(ev) - hbm0,
               note="hipMemGetInfo deltas: the FP32 handle after its first evaluation, then with the FP64 handle too")
    h32.close()
    h64.close()
    h2h = host_to_host(w) if w["kind"] == "single" else None
    if h2h:
        print(f"[bench] host to host fp32 {h2h['ms_per_step']:.2f} ms/step", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu and w["kind"] == "single":
        import d3_oracle  # noqa: F401
        cpu = cpu_baseline(w, d3_lib.table(), cn64, g64)

    sur = None if args.no_surrogate else surrogate_leg()

    out = {
        "metric": METRIC,
        "value": n / (ms32 * 1e-3),
        "unit": "atoms/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms32,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": w["desc"], "atoms": n, "functional": "PBE-D3(BJ)", "rc_2body_bohr": RC2,
                   "rc_cn_bohr": RCCN, "rc_atm_bohr": RC3, "terms": "2b+CN+ATM+chain", "precision": "fp32",
                   "table": d3_lib.table()["name"], "hip_graph": bool(args.graph), "autotuned": bool(args.autotune),
                   "bin_width_bohr": tuning32[0], "atm_blocks": tuning32[1], "atm_variant": variant32},
        "fp64_plain": {"atoms_per_s": n / (ms64 * 1e-3), "ms_per_step": ms64, "steps": k64},
        "host_to_host_fp32": h2h,
        "speedup_fp32_vs_fp64": ms64 / ms32,
        "fp32_error_vs_fp64": err,
        "phase_ms_fp32": prof32,
        "phase_ms_fp64": prof64,
        "work": {"unique_2b_pairs": uniq["two_body"], "unique_cn_pairs": uniq["cn"],
                 "unique_triangles": uniq["atm"], "flops_per_eval": total_flops,
                 "tflops_whole_step": total_flops / (ms32 * 1e-3) / 1e12},
        "roofline": {"bound": "valu", "kernel": {"atm": "k_atm<float>", "two_body": "k_2b<float>",
                                                  "cn": "k_cn<float>", "chain": "k_chain<float>"}[dom],
                     "achieved": achieved, "peak": PEAK_FP32, "unit": "TFLOP/s", "frac": achieved / PEAK_FP32,
                     "traffic": profiled_traffic({"atm": "k_atm_f32", "two_body": "k_2b_f32", "cn": "k_cn_f32",
                                                  "chain": "k_chain_f32"}[dom], w["desc"])},
        "cpu_baseline": cpu,
        "hbm_footprint": hbm,
        "surrogate": sur,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
