#!/usr/bin/env python3
"""Benchmark of the MI355X FMM evaluation pass (one JSON line on stdout).

Workload (default): 1,000,000 points uniform on the unit sphere, 128
wavelengths across the box diameter, interpolation order L = 5, leaves of at
most 64 points (Ncrit = 65), one complex right-hand side with re, im ~ U(-1, 1),
FP64 -- the reference's gemv (include/interface.hpp:153-170) as a drop-in.
A step = one evaluation pass with charges and potentials resident in HBM;
setup (octree, lists, tables, upload) is timed separately.  The same plan is
then switched to the mixed-precision path and timed again.

Accuracy: relative L2 / L-inf error on 2,000 sampled targets against an FP64
direct sum computed on the GPU.  CPU baseline: the oracle (a serial FP64
restatement of the reference, one thread) on a bounded sample of the same
family: 20,000 sphere points at the same points-per-wavelength density.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Helmholtz FMM Mpts·RHS/s at rel err≤1e-6, mixed/batched vs FP64 GPU"
PEAK_HBM_GBS = 8000.0
PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3
P2P_FLOPS_PER_PAIR = 25  # reference loop include/operators/P2P.hpp:61-75


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(n=20000, L=5, ncrit=65):
    """Oracle (port) on a sphere sample at the workload's surface density."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_api as oa
    from defmm_amd import workloads as wl
    nl = 128.0 * np.sqrt(n / 1.0e6)
    pts = wl.sphere(n, seed=1)
    kappa = wl.kappa_for(nl, pts)
    q = wl.charges(n, seed=7)
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    _, _, ts, ta = oa.fmm(pts, pts, q, L, ncrit, kappa)
    return {"value": n / ta / 1e6, "unit": "Mpts·RHS/s", "cores": 1, "kind": "port",
            "sample": f"sphere N={n} at {nl:.2f} wavelengths (same points per wavelength^2 as "
                      f"the workload), L={L}, leaf<=64, FP64, evaluation only "
                      f"({ta:.1f} s; setup {ts:.1f} s excluded)"}


def gate_regime(name, steps, warmup, samples, nrhs):
    """The same measurement, compact, on a workload where this method reaches
    relative error <= 1e-6 (order 7 at 16 wavelengths): FP64, mixed, mixed
    with nrhs right-hand sides; errors against the GPU FP64 direct sum."""
    import torch
    import defmm_amd as dm
    from defmm_amd import workloads as wl
    pts, kappa, L, ncrit = wl.make(name, seed=1)
    geo, _, nl, _, _ = wl.CONFIGS[name]
    n = len(pts)
    Q = np.stack([wl.charges(n, seed=7)] + [wl.charges(n, seed=100 + r) for r in range(1, nrhs)])
    A = dm.IBFMM_Mat(precision=dm.FP64)
    A.setSourceParticles(pts)
    A.setTargetParticles(pts)
    t0 = time.time()
    A.prcmpt(L, ncrit, kappa)
    setup_s = time.time() - t0
    idx = np.linspace(0, n - 1, samples).astype(np.int64)
    exact = dm.direct(pts, Q[0], pts[idx], kappa)
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    Qd = torch.from_numpy(Q).to(dev)
    Pd = torch.zeros((nrhs, n), dtype=torch.complex128, device=dev)

    def timed(nr):
        for _ in range(warmup):
            A.apply_batch_device(Qd.data_ptr(), Pd.data_ptr(), nr, stream.cuda_stream)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            A.apply_batch_device(Qd.data_ptr(), Pd.data_ptr(), nr, stream.cuda_stream)
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        p = Pd[0].cpu().numpy()[idx]
        err = float(np.sqrt((np.abs(p - exact) ** 2).sum() / (np.abs(exact) ** 2).sum()))
        # per-stage times of the last step, and the M2L variant that ran
        st = {k: round(v, 3) for k, v in A.stage_times().items()}
        st["m2l_variant"] = ["on the fly", "stored", "batched"][A.m2l_mode()]
        return ms, err, st

    ms64, e64, st64 = timed(1)
    A.set_precision(dm.MIXED)
    msmx, emx, stmx = timed(1)
    A.set_rhs(nrhs)
    msb, eb, stb = timed(nrhs)
    A.destroy()
    out = {"workload": f"{name}: {n:,} points ({geo}), {nl:g} wavelengths, order {L}, "
                       f"leaf<={ncrit - 1}",
           "setup_s": setup_s, "targets": samples,
           "fp64": {"value": n / (ms64 * 1e-3) / 1e6, "ms_per_step": ms64, "rel_l2_err": e64,
                    "stage_ms": st64},
           "mixed": {"value": n / (msmx * 1e-3) / 1e6, "ms_per_step": msmx, "rel_l2_err": emx,
                     "speedup_vs_fp64": ms64 / msmx, "stage_ms": stmx},
           "mixed_batched": {"value": n * nrhs / (msb * 1e-3) / 1e6, "nrhs": nrhs,
                             "ms_per_step": msb, "rel_l2_err_first_rhs": eb,
                             "speedup_vs_fp64": ms64 * nrhs / msb, "stage_ms": stb},
           "unit": "Mpts·RHS/s"}
    log(f"[gate {name}] fp64 {ms64:.2f} ms (err {e64:.3e}), mixed {msmx:.2f} ms (err {emx:.3e}), "
        f"mixed x{nrhs} {msb:.2f} ms (err {eb:.3e}); stages fp64 {st64}, mixed {stmx}, "
        f"mixed x{nrhs} {stb}")
    return out


def read_traffic(kernel, workload):
    path = os.path.join(ROOT, "profiles", f"{kernel}_traffic_{workload}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="sphere1m_128l")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--samples", type=int, default=2000)
    ap.add_argument("--m2l", default="auto", choices=["auto", "batched", "stored", "fly"],
                    help="M2L variant of the timed pass (auto: the library's choice)")
    ap.add_argument("--batch-rows", type=int, default=None,
                    help="workspace rows of the batched M2L (library default if unset)")
    ap.add_argument("--compare-m2l", action="store_true",
                    help="also time the other M2L variants (FP64)")
    ap.add_argument("--nrhs", type=int, default=2,
                    help="right-hand sides of the mixed batched pass (0 to skip); 2 keeps the "
                         "stored-Fourier M2L at 1M points / 128 wavelengths")
    ap.add_argument("--gate-nrhs", type=int, default=4,
                    help="right-hand sides of the batched pass of the 1e-6 workload")
    ap.add_argument("--gate-workload", default="sphere1m_16l_L7",
                    help="workload of the relative-error <= 1e-6 measurement ('none' to skip)")
    ap.add_argument("--copy-gib", type=float, default=0.0,
                    help="after timing, copy this many GiB device-to-device (a kernel of known "
                         "traffic, to scale rocprof byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the library has no multi-GPU path "
              f"(requested --gpus {args.gpus}); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    import defmm_amd as dm
    from defmm_amd import workloads as wl

    pts, kappa, L, ncrit = wl.make(args.workload, seed=1)
    geo, _, nl, _, _ = wl.CONFIGS[args.workload]
    n = len(pts)
    q = wl.charges(n, seed=7)
    log(f"workload {args.workload}: N={n} kappa={kappa:.6f} L={L} Ncrit={ncrit}")

    A = dm.IBFMM_Mat(precision=dm.FP64)
    A.setSourceParticles(pts)
    A.setTargetParticles(pts)
    t0 = time.time()
    A.prcmpt(L, ncrit, kappa)
    setup_s = time.time() - t0
    modes = {"batched": dm.M2L_BATCHED, "stored": dm.M2L_STORED, "fly": dm.M2L_ON_THE_FLY}
    if args.m2l == "auto":
        args.m2l = {v: k for k, v in modes.items()}[A.m2l_mode()]
    A.set_m2l(modes[args.m2l], batch_rows=args.batch_rows)
    st = A.stats()
    setup_ms = {k[3:]: st[k] for k in st if k.startswith("ms_")}
    log(f"setup {setup_s:.1f} s ({setup_ms} ms), device {A.device_bytes() / 1e9:.1f} GB, stats {st}")

    dev = torch.device("cuda:0")
    qd = torch.from_numpy(q).to(dev)
    pd = torch.zeros(n, dtype=torch.complex128, device=dev)
    stream = torch.cuda.current_stream(dev)

    # accuracy reference: direct sum on sampled targets (GPU, FP64)
    idx = np.linspace(0, n - 1, args.samples).astype(np.int64)
    exact = dm.direct(pts, q, pts[idx], kappa)

    def run(tag):
        for _ in range(args.warmup):
            A.apply_device(qd.data_ptr(), pd.data_ptr(), stream.cuda_stream)
        torch.cuda.synchronize()
        stages = []
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(args.steps):
            A.apply_device(qd.data_ptr(), pd.data_ptr(), stream.cuda_stream)
            stages.append(A.stage_times())   # events of this step (synchronises it)
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / args.steps
        p = pd.cpu().numpy()[idx]
        err2 = float(np.sqrt((np.abs(p - exact) ** 2).sum() / (np.abs(exact) ** 2).sum()))
        errinf = float(np.abs(p - exact).max() / np.abs(exact).max())
        avg = {k: float(np.mean([s[k] for s in stages])) for k in stages[0]}
        log(f"[{tag}] {ms:.2f} ms/step, rel L2 {err2:.3e}, Linf {errinf:.3e}, stages {avg}")
        return ms, err2, errinf, avg

    ms64, e64, einf64, st64 = run(f"fp64 m2l={args.m2l}")
    alt = None
    if args.compare_m2l:
        alt = {}
        for name in modes:
            if name == args.m2l:
                continue
            A.set_m2l(modes[name])
            msa, ea, _, sta = run(f"fp64 m2l={name}")
            alt[name] = {"ms_per_step": msa, "rel_l2_err": ea, "stage_ms": sta,
                         "device_gb": A.device_bytes() / 1e9}
        A.set_m2l(modes[args.m2l])
    value = n / (ms64 * 1e-3) / 1e6

    # dominant kernel of the FP64 pass and its roofline
    dom = max(["m2l_f2l", "p2p"], key=lambda k: st64[k])
    if dom == "m2l_f2l":
        # Hadamard bytes: (M2L pairs + target slots) x PP complex128, tables
        # counted once (cache-resident, negligible); the batched variant runs
        # M2F inside this stage: + F rows x (LL + PP) complex128
        PP, LL = (2 * L - 1) ** 3, L ** 3
        alg = (st["m2l_entries"] + st["m2l_targets"]) * PP * 16
        kernel = "k_m2l (M2L Hadamard + F2L)"
        if args.m2l == "batched":
            alg += st["m2l_batch_rows"] * (LL + PP) * 16
            kernel = (f"k_m2f + k_m2l per batch ({st['m2l_batches']} batches of <= "
                      f"{args.batch_rows or 16384} F rows): M2F, M2L Hadamard, F2L")
        achieved = alg / (st64[dom] * 1e-3) / 1e9
        roof = {"kernel": kernel, "bound": "hbm", "achieved": achieved,
                "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS,
                "traffic": read_traffic(f"m2l_{args.m2l}", args.workload),
                "algorithmic_bytes": alg, "avg_ms": st64[dom]}
    else:
        flops = st["p2p_pairs"] * P2P_FLOPS_PER_PAIR
        achieved = flops / (st64[dom] * 1e-3) / 1e12
        roof = {"kernel": "k_p2p (+k_p2p_reduce)", "bound": "valu", "achieved": achieved,
                "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS,
                "traffic": None, "algorithmic_flops": flops, "avg_ms": st64[dom]}

    mixed = None
    if not args.no_mixed:
        t0 = time.time()
        A.set_precision(dm.MIXED)
        log(f"switch to mixed: {time.time() - t0:.1f} s, device {A.device_bytes() / 1e9:.1f} GB")
        msmx, emx, einfmx, stmx = run("mixed")
        mixed = {"value": n / (msmx * 1e-3) / 1e6, "ms_per_step": msmx, "rel_l2_err": emx,
                 "rel_linf_err": einfmx, "speedup_vs_fp64": ms64 / msmx, "stage_ms": stmx,
                 "dtype": "f32 storage, f64 phase reduction"}
    batched = None
    if not args.no_mixed and args.nrhs > 1:
        nr = args.nrhs
        t0 = time.time()
        A.set_rhs(nr)
        log(f"width {nr} right-hand sides: {time.time() - t0:.1f} s, device "
            f"{A.device_bytes() / 1e9:.1f} GB, M2L variant {A.m2l_mode()}")
        Q = np.stack([q] + [wl.charges(n, seed=100 + r) for r in range(1, nr)])
        Qd = torch.from_numpy(Q).to(dev)
        Pd = torch.zeros((nr, n), dtype=torch.complex128, device=dev)
        for _ in range(args.warmup):
            A.apply_batch_device(Qd.data_ptr(), Pd.data_ptr(), nr, stream.cuda_stream)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        stages = []
        e0.record(stream)
        for _ in range(args.steps):
            A.apply_batch_device(Qd.data_ptr(), Pd.data_ptr(), nr, stream.cuda_stream)
            stages.append(A.stage_times())
        e1.record(stream)
        torch.cuda.synchronize()
        msb = e0.elapsed_time(e1) / args.steps
        p0 = Pd[0].cpu().numpy()[idx]
        eb = float(np.sqrt((np.abs(p0 - exact) ** 2).sum() / (np.abs(exact) ** 2).sum()))
        avg = {k: float(np.mean([s[k] for s in stages])) for k in stages[0]}
        log(f"[mixed x{nr}] {msb:.2f} ms/pass, {msb / nr:.2f} ms per right-hand side, rel L2 "
            f"(first) {eb:.3e}, stages {avg}")
        batched = {"value": n * nr / (msb * 1e-3) / 1e6, "nrhs": nr, "ms_per_step": msb,
                   "rel_l2_err_first_rhs": eb, "speedup_vs_fp64": ms64 * nr / msb,
                   "stage_ms": avg, "m2l_variant": ["on the fly", "stored", "batched"][A.m2l_mode()],
                   "device_gb": A.device_bytes() / 1e9,
                   "dtype": "f32 storage, f64 phase reduction"}
        del Qd, Pd
    A.destroy()
    if args.copy_gib > 0:
        nb = int(args.copy_gib * (1 << 30)) // 16
        src = torch.ones(nb, dtype=torch.complex128, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copied {nb * 16} bytes device-to-device")
        del src, dst

    gate = None
    if args.gate_workload != "none":
        gate = gate_regime(args.gate_workload, args.steps, args.warmup, args.samples,
                           args.gate_nrhs)

    cpu = None
    if not args.no_cpu:
        t0 = time.time()
        cpu = cpu_baseline()
        log(f"cpu baseline {cpu} ({time.time() - t0:.1f} s)")

    out = {
        "metric": METRIC, "value": value, "unit": "Mpts·RHS/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms64,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {n:,} points ({geo}), {nl:g} wavelengths, "
                               f"order {L}, leaf<={ncrit - 1}, 1 RHS, FP64 drop-in",
                   "n_points": n, "kappa": kappa, "order": L, "ncrit": ncrit, "nrhs": 1,
                   "parallelism": "single GPU"},
        "accuracy": {"rel_l2_err": e64, "rel_linf_err": einf64, "targets": args.samples,
                     "reference": "FP64 direct sum (GPU) on sampled targets"},
        "setup_s": setup_s,
        "setup_ms_by_phase": setup_ms,
        "stage_ms": st64,
        "roofline": roof,
        "cpu_baseline": cpu,
        "mixed": mixed,
        "mixed_batched": batched,
        "accuracy_gate_1e-6": gate,
        "m2l_variants_fp64": alt,
        "plan": {k: st[k] for k in ("p2p_pairs", "hf_m2l", "src_slots", "trg_slots",
                                      "m2f_slots", "m2l_entries", "m2l_group_entries",
                                      "m2l_batches", "m2l_batch_rows", "p2p_items")},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
