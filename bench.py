#!/usr/bin/env python3
"""Benchmark of the MI355X FMM evaluation pass (one JSON line on stdout).

Metric: Helmholtz FMM throughput in Mpts*RHS/s (targets x right-hand sides
per second of evaluation) at relative error <= 1e-6 against the FP64 direct
sum, mixed-precision batched path against the plain FP64 path.

Headline workload (default): 20,000 points on the surface of the unit cube,
refined toward the edges, 32 wavelengths across the box, leaves of at most 64
points, complex charges re, im ~ U(-1, 1), at interpolation order 8 -- the
lowest order whose error is <= 1e-6 on this input (7.2e-7; order 7 gives
5.8e-6, order 4 of the workload's nominal setting 1.1e-3).  A step is one
evaluation pass (the reference's gemv, include/interface.hpp:153-170) with
charges and potentials resident in HBM; setup (octree, lists, tables,
upload) is timed separately.  Measured on one plan:
  fp64           FP64, one right-hand side (the drop-in)
  mixed          float expansions and near field, double phase reduction
  mixed_batched  the same with --nrhs right-hand sides per pass: `value`
Errors: relative L2 / L-inf on --samples sampled targets against an FP64
direct sum computed on the GPU.

Instance batch (--batch-instances): the uneven instance mix of the
benchmark's instance-batch configuration at reduced size (48 instances of
10,000 - 100,000 points instead of 256 of 10,000 - 500,000, which would not
fit one GPU's memory at once): each instance evaluated by its own launches
one after another, then all of them by one HIP graph (defmm_batch_*).

Scale section (--scale-workload, default the 1M-point sphere at 128
wavelengths, order 5): FP64 and mixed at full size; at this frequency order 5
gives 7e-3, and the order that reaches 1e-6 there (about 10) does not fit in
one GPU's memory, so it reports throughput and error as measured.  It
also times the same plan's set-up by the host builder alone (host-only
build, no upload) against the default device set-up without its upload.

Roofline: the dominant stage of the value pass, the fused M2L+F2L (with, in
the batched M2L variant, the M2F launches that refill its workspace).
Compulsory HBM bytes = source expansions (batched: M rows; stored: Fourier
rows) read once per right-hand side + kernel table rows read once per
launch + local expansions written once per right-hand side + the group
lists; divided by the stage's time measured with HIP events on the launch
stream.  `traffic` is the stage's FETCH_SIZE + WRITE_SIZE per pass from
rocprofv3 counter runs of the same variant (profiles/m2l_traffic_*.txt,
written by profile_summary.py), scaled by a copy kernel of known size.

CPU baseline: the oracle (serial FP64 restatement of the reference) at order
8 on a reduced twin of the workload (8,000 cube points at the same points per
wavelength squared), timed with one thread and with one evaluation per core
running concurrently (the reference is serial and holds one instance per
process).
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


def m2l_nr_per_launch(L, mixed):
    """Right-hand sides one M2L launch serves (the executor's rule: at most 16
    double or 32 float complex accumulators per thread; 1 in double from
    order 7, 2 in float at order 7; from order 8 a thread holds a column of
    2L - 1 frequencies: 1 in double, 2 in float)."""
    PP = (2 * L - 1) ** 3
    if L >= 8:  # one column of 2L - 1 frequencies per thread
        return 2 if mixed else 1
    nw = max(-(-PP // 1024), min(-(-PP // 256), 3))
    g = 3 if L == 7 else 4
    cap = (32 if mixed else 16) // (g * nw)
    w = 8 if cap >= 8 else 4 if cap >= 4 else 2 if cap >= 2 else 1
    if not mixed and L >= 7:
        return 1
    if L == 7 and w > 2:
        return 2
    return w


def m2l_compulsory_bytes(st, L, nrhs, mixed, batched=False):
    """Compulsory HBM bytes of the M2L+F2L stage for nrhs right-hand sides:
    stored variant: every F row read once per right-hand side; batched
    variant: every source M row read once per right-hand side (its F row is
    made by the batch's M2F and read back while cache-resident); both: every
    table row once per launch over all groups, every local expansion written
    once per right-hand side, and the group lists once per launch."""
    PP, LL = (2 * L - 1) ** 3, L ** 3
    s = 8 if mixed else 16
    g = 1 if L >= 8 else 3 if L == 7 else 4
    nr = m2l_nr_per_launch(L, mixed)
    launches = -(-nrhs // nr)
    lists = st["m2l_group_entries"] * (4 + 4 + 4 * g) + st["m2l_groups"] * (4 + 4 * g)
    per_rhs = st["m2f_slots"] * (LL if batched else PP) * s + st["m2l_targets"] * LL * s
    per_launch = st["m2l_table_rows"] * PP * s + lists
    return nrhs * per_rhs + launches * per_launch, launches


def cpu_info():
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count()
    return model, os.cpu_count(), avail


def cpu_baseline(L=8, n=8000, cores=None):
    """Oracle (port) at order L on a reduced twin of the headline workload:
    n cube points at the same points per wavelength squared; one thread, then
    `cores` concurrent evaluations (different charges), apply time only."""
    from concurrent.futures import ThreadPoolExecutor
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_api as oa
    from defmm_amd import workloads as wl
    model, nproc, avail = cpu_info()
    if cores is None:
        cores = min(16, avail)
    nl = 32.0 * np.sqrt(n / 20000.0)
    pts = wl.cube_edges(n, seed=1)
    kappa = wl.kappa_for(nl, pts)
    _, st, ts, ta1 = oa.fmm(pts, pts, wl.charges(n, seed=7), L, 65, kappa)
    one = n / ta1 / 1e6

    def job(r):
        return oa.fmm(pts, pts, wl.charges(n, seed=900 + r), L, 65, kappa)[3]

    t0 = time.time()
    with ThreadPoolExecutor(cores) as ex:
        tas = list(ex.map(job, range(cores)))
    wall = time.time() - t0
    agg = sum(n / t / 1e6 for t in tas)
    return {"value": agg, "unit": "Mpts·RHS/s", "cores": cores, "kind": "port",
            "value_1_thread": one, "cpu_model": model, "nproc": nproc,
            "cores_available": avail,
            "sample": f"oracle at order {L} on {n:,} edge-refined cube points at {nl:.2f} "
                      f"wavelengths (the workload's points per wavelength^2), leaf<=64, FP64, "
                      f"evaluation only: 1 thread {ta1:.1f} s; {cores} concurrent evaluations "
                      f"(one per core, different charges) {np.mean(tas):.1f} s each, "
                      f"{wall:.1f} s wall; setup {ts:.1f} s excluded"}


class Runner:
    def __init__(self, name, samples, steps, warmup):
        import torch
        import defmm_amd as dm
        from defmm_amd import workloads as wl
        self.torch, self.dm, self.wl = torch, dm, wl
        self.name, self.steps, self.warmup = name, steps, warmup
        self.pts, self.kappa, self.L, self.ncrit = wl.make(name, seed=1)
        geo, _, self.nl, _, _ = wl.CONFIGS[name]
        self.geo = geo
        self.n = len(self.pts)
        self.q = wl.charges(self.n, seed=7)
        log(f"workload {name}: N={self.n} kappa={self.kappa:.6f} L={self.L} Ncrit={self.ncrit}")
        A = dm.IBFMM_Mat(precision=dm.FP64)
        A.setSourceParticles(self.pts)
        A.setTargetParticles(self.pts)
        t0 = time.time()
        A.prcmpt(self.L, self.ncrit, self.kappa)
        self.setup_s = time.time() - t0
        self.A = A
        self.st = A.stats()
        self.setup_ms = {k[3:]: self.st[k] for k in self.st if k.startswith("ms_")}
        log(f"setup {self.setup_s:.1f} s ({self.setup_ms} ms), device "
            f"{A.device_bytes() / 1e9:.1f} GB, M2L variant {A.m2l_mode()}")
        self.dev = torch.device("cuda:0")
        self.stream = torch.cuda.current_stream(self.dev)
        self.idx = np.linspace(0, self.n - 1, samples).astype(np.int64)
        self.exact = dm.direct(self.pts, self.q, self.pts[self.idx], self.kappa)

    def describe(self):
        return (f"{self.name}: {self.n:,} points ({self.geo}), {self.nl:g} wavelengths, "
                f"order {self.L}, leaf<={self.ncrit - 1}")

    def timed(self, nrhs, tag):
        """Time `steps` passes of nrhs right-hand sides (device arrays);
        returns ms per pass, errors of the first right-hand side, mean stage
        times."""
        torch, A = self.torch, self.A
        Q = np.stack([self.q] + [self.wl.charges(self.n, seed=100 + r) for r in range(1, nrhs)])
        Qd = torch.from_numpy(Q).to(self.dev)
        Pd = torch.zeros((nrhs, self.n), dtype=torch.complex128, device=self.dev)
        s = self.stream.cuda_stream
        for _ in range(self.warmup):
            A.apply_batch_device(Qd.data_ptr(), Pd.data_ptr(), nrhs, s)
        torch.cuda.synchronize()
        stages = []
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(self.stream)
        for _ in range(self.steps):
            A.apply_batch_device(Qd.data_ptr(), Pd.data_ptr(), nrhs, s)
            stages.append(A.stage_times())   # events of this step (synchronises it)
        e1.record(self.stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / self.steps
        p = Pd[0].cpu().numpy()[self.idx]
        ex = self.exact
        e2 = float(np.sqrt((np.abs(p - ex) ** 2).sum() / (np.abs(ex) ** 2).sum()))
        einf = float(np.abs(p - ex).max() / np.abs(ex).max())
        avg = {k: float(np.mean([x[k] for x in stages])) for k in stages[0]}
        log(f"[{self.name} {tag}] {ms:.3f} ms/pass ({nrhs} rhs), rel L2 {e2:.3e}, Linf "
            f"{einf:.3e}, stages {({k: round(v, 3) for k, v in avg.items()})}")
        del Qd, Pd
        return ms, e2, einf, avg

    def roofline(self, stage_ms, nrhs, mixed):
        """Roofline of the dominant kernel of a pass (M2L+F2L or P2P)."""
        st, L = self.A.stats(), self.L  # (batch counts change with the M2L variant)
        dom = max(["m2l_f2l", "p2p"], key=lambda k: stage_ms[k])
        t = stage_ms[dom] * 1e-3
        mode = self.A.m2l_mode()
        if dom == "m2l_f2l" and mode in (self.dm.M2L_STORED, self.dm.M2L_BATCHED):
            batched = mode == self.dm.M2L_BATCHED
            comp, launches = m2l_compulsory_bytes(st, L, nrhs, mixed, batched)
            achieved = comp / t / 1e9
            kname = "k_m2l_col" if L >= 8 else "k_m2l"
            # the bytes the Hadamard reads from L2 / MALL / HBM: an F row and
            # a table row per pair (the table row once per launch's right-hand
            # sides)
            s_c = 8 if mixed else 16
            nr = m2l_nr_per_launch(L, mixed)
            pair_bytes = st["m2l_entries"] * (2 * L - 1) ** 3 * s_c * (nrhs + -(-nrhs // nr))
            out = {"kernel": f"{kname} (fused M2L Hadamard + F2L)" +
                             (f", with k_m2f per batch of F rows ({st['m2l_batches']} batches)"
                              if batched else ""),
                   "m2l_variant": "batched" if batched else "stored",
                   "bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                   "frac": achieved / PEAK_HBM_GBS, "traffic": None,
                   "compulsory_bytes_per_pass": comp, "launches": launches,
                   "avg_ms_per_launch": stage_ms[dom] / launches,
                   "pair_bytes_read_GBs": pair_bytes / t / 1e9,
                   "dtype": "f32" if mixed else "f64"}
            if not batched:
                out["compulsory_bytes_per_launch"] = comp / launches
            return out
        flops = st["p2p_pairs"] * P2P_FLOPS_PER_PAIR * nrhs
        peak = PEAK_FP32_TFLOPS if mixed else PEAK_FP64_TFLOPS
        achieved = flops / t / 1e12
        return {"kernel": "k_p2p + k_p2p_reduce", "bound": "valu", "achieved": achieved,
                "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak, "traffic": None,
                "algorithmic_flops": flops, "avg_ms": stage_ms[dom]}

    def close(self):
        self.A.destroy()


def batch_instances(count, nmin, nmax, seed=5150):
    """The uneven instance mix of the instance-batch configuration: sizes
    log-uniform in [nmin, nmax], geometry uniform over {sphere, edge-refined
    cube, clustered ellipsoid, 16 Gaussian clusters}, 2^U(3,7) wavelengths,
    order 4-6, 1-16 right-hand sides; one seed per instance."""
    from defmm_amd import workloads as wl
    u = wl.uniform(seed, 8 * count)
    geos = ["sphere", "cube_edges", "ellipsoid", "clusters"]
    out = []
    for i in range(count):
        n = int(10 ** (np.log10(nmin) + (np.log10(nmax) - np.log10(nmin)) * u[8 * i]))
        geo = geos[min(3, int(u[8 * i + 1] * 4))]
        nl = 2 ** (3 + 4 * u[8 * i + 2])
        L = 4 + min(2, int(u[8 * i + 3] * 3))
        nr = 1 + min(15, int(u[8 * i + 4] * 16))
        out.append((i, geo, n, nl, L, nr))
    return out


def batch_section(count, nmin, nmax, width, steps, warmup):
    """Instances evaluated one after another (each plan's own launches) and
    as one graph (defmm_batch_*); mixed precision; plan width <= `width`
    right-hand sides (wider instances run in chunks of it)."""
    import torch
    import defmm_amd as dm
    from defmm_amd import workloads as wl
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    mats, Qs, Ps, meta = [], [], [], []
    t0 = time.time()
    skipped = 0
    for (i, geo, n, nl, L, nr) in batch_instances(count, nmin, nmax):
        pts = wl.GEOMETRIES[geo](n, seed=900 + i)
        kappa = wl.kappa_for(nl, pts)
        A = dm.IBFMM_Mat(precision=dm.MIXED)
        A.setSourceParticles(pts)
        A.setTargetParticles(pts)
        try:
            A.prcmpt(L, 65, kappa)
            A.set_rhs(min(nr, width))
        except dm.DefmmError as e:
            log(f"[batch] instance {i} ({geo} N={n} {nl:.1f} wavelengths L={L}): {e}; left out")
            A.destroy()
            skipped += 1
            continue
        Q = np.stack([wl.charges(n, seed=1000 + 17 * i + r) for r in range(nr)])
        mats.append(A)
        meta.append((i, geo, n, nl, L, nr, pts, kappa, Q[0]))
        Qs.append(torch.from_numpy(Q).to(dev))
        Ps.append(torch.zeros((nr, n), dtype=torch.complex128, device=dev))
    setup_s = time.time() - t0
    work = sum(m[2] * m[5] for m in meta)
    dev_gb = sum(A.device_bytes() for A in mats) / 1e9
    log(f"[batch] {len(mats)} instances ({skipped} left out), {sum(m[2] for m in meta):,} points, "
        f"{work:,} point-RHS, setup {setup_s:.1f} s, device {dev_gb:.1f} GB")

    def seq():
        for A, m, q, p in zip(mats, meta, Qs, Ps):
            A.apply_batch_device(q.data_ptr(), p.data_ptr(), m[5], stream.cuda_stream)

    def timeit(fn):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps

    ms_seq = timeit(seq)
    ref = [p.clone() for p in Ps]
    t0 = time.time()
    B = dm.Batch(mats, [q.data_ptr() for q in Qs], [p.data_ptr() for p in Ps],
                 [m[5] for m in meta], nbranch=8)
    capture_s = time.time() - t0
    ms_graph = timeit(lambda: B.execute(stream.cuda_stream))
    same = all(torch.equal(a, b) for a, b in zip(ref, Ps))
    # accuracy of the first right-hand side on 200 sampled targets of every
    # fourth instance, against the FP64 direct sum
    errs = []
    for k in range(0, len(meta), 4):
        i, geo, n, nl, L, nr, pts, kappa, q0 = meta[k]
        idx = np.linspace(0, n - 1, 200).astype(np.int64)
        ex = dm.direct(pts, q0, pts[idx], kappa)
        p = Ps[k][0].cpu().numpy()[idx]
        errs.append((i, L, float(np.sqrt((np.abs(p - ex) ** 2).sum() / (np.abs(ex) ** 2).sum()))))
    B.destroy()
    for A in mats:
        A.destroy()
    log(f"[batch] sequential {ms_seq:.2f} ms, graph {ms_graph:.2f} ms (capture {capture_s:.2f} s), "
        f"identical {same}, errors {errs}")
    return {"workload": f"{len(mats)} instances, {nmin:,}-{nmax:,} points (log-uniform), "
                        f"sphere / edge-refined cube / clustered ellipsoid / 16 Gaussian clusters, "
                        f"8-128 wavelengths, order 4-6, 1-16 RHS, mixed precision, leaf<=64",
            "instances": len(mats), "left_out": skipped, "points": sum(m[2] for m in meta),
            "point_rhs": work, "setup_s": setup_s, "device_gb": dev_gb,
            "sequential": {"ms_per_step": ms_seq, "value": work / (ms_seq * 1e-3) / 1e6},
            "graph": {"ms_per_step": ms_graph, "value": work / (ms_graph * 1e-3) / 1e6,
                      "capture_s": capture_s, "identical_to_sequential": same},
            "rel_l2_err_sampled": errs, "unit": "Mpts·RHS/s"}


def multi_rhs_section(name, nrhs, width, steps, warmup, samples=200, autotune=True):
    """A many-right-hand-side workload at full size: FP64 with one
    right-hand side, then mixed precision with nrhs right-hand sides evaluated
    `width` at a time (each after the autotuner, one timed pass per
    candidate); errors of the first right-hand side on sampled targets
    against the FP64 direct sum."""
    import torch
    import defmm_amd as dm
    R = Runner(name, samples, steps, warmup)
    tunes = {}
    if autotune:
        t0 = time.time()
        tunes["fp64"] = R.A.autotune(1, reps=1)
        log(f"autotune {name} fp64 ({time.time() - t0:.1f} s):\n{tunes['fp64']}")
    ms64, e64, _, st64 = R.timed(1, "fp64")
    R.A.set_precision(dm.MIXED)
    # (the FP64 tuning fixed the M2L variant for one right-hand side; the
    # batched variant fits any width, and the tuning below chooses again)
    R.A.set_m2l(dm.M2L_BATCHED)
    R.A.set_rhs(width)
    if autotune:
        t0 = time.time()
        tunes["mixed_batched"] = R.A.autotune(width, reps=1)
        log(f"autotune {name} mixed x{width} ({time.time() - t0:.1f} s):\n"
            f"{tunes['mixed_batched']}")
    msb, eb, _, stb = R.timed(nrhs, f"mixed x{nrhs} (width {width})")
    out = {"workload": R.describe(), "setup_s": R.setup_s, "setup_ms_by_phase": R.setup_ms,
           "plan": {k: R.st[k] for k in ("p2p_pairs", "hf_m2l", "src_slots", "trg_slots",
                                         "m2f_slots", "m2l_table_rows")},
           "fp64": {"value": R.n / (ms64 * 1e-3) / 1e6, "ms_per_step": ms64, "rel_l2_err": e64,
                    "stage_ms": st64, "autotune": tunes.get("fp64")},
           "mixed_batched": {"value": R.n * nrhs / (msb * 1e-3) / 1e6, "nrhs": nrhs,
                             "width": width, "ms_per_step": msb, "rel_l2_err_first_rhs": eb,
                             "stage_ms": stb, "speedup_vs_fp64": ms64 * nrhs / msb,
                             "m2l_variant": ["on the fly", "stored", "batched"][R.A.m2l_mode()],
                             "device_gb": R.A.device_bytes() / 1e9,
                             "autotune": tunes.get("mixed_batched")},
           "unit": "Mpts·RHS/s"}
    R.close()
    torch.cuda.empty_cache()
    return out


def read_traffic(kernel, workload, key="traffic_bytes_per_launch"):
    """A measured traffic figure from a profile summary under profiles/."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_traffic_{workload}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key + ":"):
            return float(line.split(":")[1])
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="cube20k_32l_L8")
    ap.add_argument("--nrhs", type=int, default=16,
                    help="right-hand sides of the mixed batched pass (the value)")
    ap.add_argument("--samples", type=int, default=1000)
    ap.add_argument("--scale-workload", default="sphere1m_128l",
                    help="full-size workload timed in FP64 and mixed ('none' to skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="time every pass with the library's default settings")
    ap.add_argument("--m2l-rows", type=int, default=0,
                    help="time the mixed batched pass with the batched M2L variant and this many "
                         "workspace rows instead of autotuning it (profiling runs)")
    ap.add_argument("--batch-instances", type=int, default=48,
                    help="instances of the uneven instance batch (0 to skip)")
    ap.add_argument("--batch-sizes", default="10000,100000",
                    help="smallest,largest instance size of the instance batch")
    ap.add_argument("--multi-rhs-workload", default="none",
                    help="full-size many-right-hand-side workload, e.g. ellipsoid2m_256l_L6 "
                         "('none' to skip)")
    ap.add_argument("--multi-rhs", type=int, default=64)
    ap.add_argument("--multi-rhs-width", type=int, default=16)
    ap.add_argument("--batch-width", type=int, default=2,
                    help="plan width (right-hand sides held at once) of the batch's instances")
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

    R = Runner(args.workload, args.samples, args.steps, args.warmup)
    tunes = {}

    def tune(tag, nr):
        if args.no_autotune:
            return
        t0 = time.time()
        tunes[tag] = R.A.autotune(nr, reps=3)
        log(f"autotune {tag} ({time.time() - t0:.1f} s):\n{tunes[tag]}")

    tune("fp64", 1)
    ms64, e64, einf64, st64 = R.timed(1, "fp64")
    roof64 = R.roofline(st64, 1, False)
    v64_variant = ["on the fly", "stored", "batched"][R.A.m2l_mode()]
    R.A.set_precision(dm.MIXED)
    tune("mixed", 1)
    msmx, emx, einfmx, stmx = R.timed(1, "mixed")
    R.A.set_rhs(args.nrhs)
    if args.m2l_rows > 0:
        R.A.set_m2l(dm.M2L_BATCHED, batch_rows=args.m2l_rows)
    else:
        tune("mixed_batched", args.nrhs)
    msb, eb, einfb, stb = R.timed(args.nrhs, f"mixed x{args.nrhs}")
    roof = R.roofline(stb, args.nrhs, True)
    if roof.get("m2l_variant"):
        # HBM bytes of the M2L stage per pass, measured with rocprofv3 counters
        # in the same variant (profiles/m2l_traffic_<workload>.txt)
        roof["traffic"] = read_traffic("m2l", args.workload,
                                       f"{roof['m2l_variant']}_f32_x{args.nrhs}_bytes_per_pass")
        if roof["traffic"]:
            roof["traffic_over_compulsory"] = roof["traffic"] / roof["compulsory_bytes_per_pass"]
    variant = ["on the fly", "stored", "batched"][R.A.m2l_mode()]
    n = R.n
    v64 = n / (ms64 * 1e-3) / 1e6
    vmx = n / (msmx * 1e-3) / 1e6
    value = n * args.nrhs / (msb * 1e-3) / 1e6
    head = {"workload": R.describe(), "setup_s": R.setup_s, "setup_ms_by_phase": R.setup_ms,
            "plan": {k: R.st[k] for k in ("p2p_pairs", "hf_m2l", "src_slots", "trg_slots",
                                          "m2f_slots", "m2l_entries", "m2l_group_entries",
                                          "m2l_table_rows", "p2p_items")}}
    R.close()
    if args.copy_gib > 0:
        nb = int(args.copy_gib * (1 << 30)) // 16
        src = torch.ones(nb, dtype=torch.complex128, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copied {nb * 16} bytes device-to-device")
        del src, dst

    scale = None
    if args.scale_workload != "none":
        S = Runner(args.scale_workload, 2000, max(1, min(args.steps, 5)), 1)
        stunes = {}
        if not args.no_autotune:
            t0 = time.time()
            stunes["fp64"] = S.A.autotune(1, reps=1)
            log(f"autotune scale fp64 ({time.time() - t0:.1f} s):\n{stunes['fp64']}")
        s64, se64, _, sst64 = S.timed(1, "fp64")
        sroof = S.roofline(sst64, 1, False)
        if sroof.get("m2l_variant"):
            sroof["traffic"] = read_traffic("m2l", args.scale_workload,
                                            f"{sroof['m2l_variant']}_f64_x1_bytes_per_pass")
            if sroof["traffic"]:
                sroof["traffic_over_compulsory"] = (sroof["traffic"] /
                                                    sroof["compulsory_bytes_per_pass"])
        S.A.set_precision(dm.MIXED)
        if not args.no_autotune:
            t0 = time.time()
            stunes["mixed"] = S.A.autotune(1, reps=1)
            log(f"autotune scale mixed ({time.time() - t0:.1f} s):\n{stunes['mixed']}")
        smx, semx, _, sstmx = S.timed(1, "mixed")
        scale = {"workload": S.describe(), "setup_s": S.setup_s,
                 "setup_ms_by_phase": S.setup_ms, "device_gb_mixed": S.A.device_bytes() / 1e9,
                 "fp64": {"value": S.n / (s64 * 1e-3) / 1e6, "ms_per_step": s64,
                          "rel_l2_err": se64, "stage_ms": sst64, "roofline": sroof,
                          "autotune": stunes.get("fp64")},
                 "mixed": {"value": S.n / (smx * 1e-3) / 1e6, "ms_per_step": smx,
                           "rel_l2_err": semx, "stage_ms": sstmx,
                           "speedup_vs_fp64": s64 / smx, "autotune": stunes.get("mixed")},
                 "note": "order 5 at 128 wavelengths: error 7e-3, above the 1e-6 gate"}
        S.close()
        # the same plan built by the host builder alone (no device work, no
        # upload), against the device set-up without its upload
        t0 = time.time()
        hs = dm.build_plan_stats(S.pts, S.pts, S.L, S.ncrit, S.kappa)
        host_s = time.time() - t0
        dev_s = S.setup_s - S.setup_ms.get("upload", 0) / 1e3
        scale["setup_host_builder_s"] = host_s
        scale["setup_device_builder_s_without_upload"] = dev_s
        scale["setup_plans_equal"] = all(hs[k] == S.st[k] for k in hs
                                         if not k.startswith("ms_") and k not in
                                         ("setup_device", "m2l_batches", "m2l_batch_rows",
                                          "m2l_table_rows", "p2p_items"))
        log(f"set-up without upload: device builder {dev_s:.2f} s, host builder {host_s:.2f} s, "
            f"same lists {scale['setup_plans_equal']}")

    many = None
    if args.multi_rhs_workload != "none":
        many = multi_rhs_section(args.multi_rhs_workload, args.multi_rhs, args.multi_rhs_width,
                                 max(1, min(args.steps, 3)), 1, autotune=not args.no_autotune)

    batch = None
    if args.batch_instances > 0:
        lo, hi = (int(x) for x in args.batch_sizes.split(","))
        batch = batch_section(args.batch_instances, lo, hi, args.batch_width,
                              max(1, min(args.steps, 5)), 1)

    cpu = None
    if not args.no_cpu:
        t0 = time.time()
        cpu = cpu_baseline(L=R.L)
        log(f"cpu baseline {cpu} ({time.time() - t0:.1f} s)")

    gate_ok = max(e64, emx, eb) <= 1e-6
    out = {
        "metric": METRIC, "value": value, "unit": "Mpts·RHS/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": msb,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{R.describe()}, {args.nrhs} RHS, mixed precision (float "
                               f"expansions, double phase reduction), {variant} M2L",
                   "n_points": n, "kappa": R.kappa, "order": R.L, "ncrit": R.ncrit,
                   "nrhs": args.nrhs, "parallelism": "single GPU"},
        "accuracy": {"rel_l2_err": eb, "rel_linf_err": einfb, "targets": args.samples,
                     "gate_1e-6_met": gate_ok,
                     "reference": "FP64 direct sum (GPU) on sampled targets"},
        "fp64": {"value": v64, "ms_per_step": ms64, "rel_l2_err": e64, "rel_linf_err": einf64,
                 "stage_ms": st64, "roofline": roof64, "m2l_variant": v64_variant,
                 "autotune": tunes.get("fp64")},
        "mixed": {"value": vmx, "ms_per_step": msmx, "rel_l2_err": emx, "rel_linf_err": einfmx,
                  "speedup_vs_fp64": v64 and vmx / v64, "stage_ms": stmx,
                  "autotune": tunes.get("mixed")},
        "mixed_batched": {"value": value, "nrhs": args.nrhs, "ms_per_step": msb,
                          "rel_l2_err_first_rhs": eb, "speedup_vs_fp64": value / v64,
                          "stage_ms": stb, "m2l_variant": variant,
                          "autotune": tunes.get("mixed_batched")},
        "speedup_mixed_batched_vs_fp64": value / v64,
        "roofline": roof,
        "cpu_baseline": cpu,
        "headline": head,
        "scale": scale,
        "instance_batch": batch,
        "many_rhs": many,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
