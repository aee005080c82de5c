#!/usr/bin/env python3
"""Reduce rocprofv3 CSV output of bench.py runs to the text summaries kept
under profiles/.

  kernels  DIR  --out FILE --cmd CMD
      per-kernel duration statistics (kernel_stats.csv) and, per evaluation
      pass found in kernel_trace.csv, the M2L stage's kernels
  traffic  FETCH_DIR WRITE_DIR --before W1 [--after W2] --cmd CMD --cmd CMD
      HBM bytes per pass of the M2L stage from a FETCH_SIZE run and a
      WRITE_SIZE run of the same command, scaled by the device-to-device copy
      that bench.py --copy-gib 4 makes between its headline and scale
      sections; passes before the copy belong to workload W1, after it to W2;
      writes profiles/m2l_traffic_<W>.txt with the keys bench.py reads
  counters DIR --out FILE --cmd CMD
      per-kernel instruction counters (VALU, MFMA) and matrix-core busy
      fraction of one --pmc run

A pass is the run of dispatches from one k_zero of the potentials (the
executor's first launch of a pass) to the next.  Its M2L stage is every
k_m2l* launch, plus its k_m2f launches when the batched variant interleaves
them with the M2L launches (a k_m2f after the first k_m2l of the pass).
"""
import argparse
import collections
import csv
import glob
import os
import re
import sys

ROOT = os.path.dirname(os.path.abspath(__file__))
COPY_BYTES = 4 << 30


def short(name):
    name = name.replace("HIP_vector_type<double, 2u>", "double2")
    name = name.replace("HIP_vector_type<float, 2u>", "float2")
    name = re.sub(r"^void ", "", name)
    name = name.replace("(anonymous namespace)::", "")
    m = re.match(r"([\w:]+)(<[^(]*>)?\(", name)
    return (m.group(1) + (m.group(2) or "")) if m else name[:70]


def one(dirpath, suffix):
    hits = sorted(glob.glob(os.path.join(dirpath, "**", "*" + suffix), recursive=True))
    if not hits:
        sys.exit(f"no *{suffix} under {dirpath}")
    return hits[0]


def rows(path):
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def dispatches(path, counter=None):
    """[(dispatch id, short kernel name, value or duration ns, grid)] in order."""
    out = []
    for r in rows(path):
        if counter is not None:
            if r["Counter_Name"] != counter:
                continue
            v = float(r["Counter_Value"])
        else:
            v = float(r["End_Timestamp"]) - float(r["Start_Timestamp"])
        grid = r.get("Grid_Size") or r.get("Grid_Size_X") or ""
        out.append((int(r["Dispatch_Id"]), short(r["Kernel_Name"]), v, grid))
    out.sort()
    return out


def is_m2l(k):
    return k.startswith("dfm::k_m2l")


def passes(ds):
    """Split dispatches into evaluation passes; returns (start index, list)."""
    cuts = [i for i, d in enumerate(ds) if d[1].startswith("dfm::k_zero<")]
    out = []
    for a, b in zip(cuts, cuts[1:] + [len(ds)]):
        seg = ds[a:b]
        # a pass ends with its scatters: drop whatever follows the last one
        last = max((i for i, d in enumerate(seg) if d[1].startswith("dfm::k_scatter")),
                   default=len(seg) - 1)
        out.append((a, seg[:last + 1]))
    return out


def describe(seg):
    names = [d[1] for d in seg]
    m2l = [i for i, k in enumerate(names) if is_m2l(k)]
    prec = "f32" if any("<float" in k for k in names if is_m2l(k)) else "f64"
    nrhs = sum(1 for k in names if k.startswith("dfm::k_gather"))
    if not m2l:
        return prec, nrhs, "none", []
    first = m2l[0]
    batched = any(k.startswith("dfm::k_m2f") for k in names[first:])
    stage = [i for i, k in enumerate(names)
             if is_m2l(k) or (batched and k.startswith("dfm::k_m2f"))]
    return prec, nrhs, "batched" if batched else "stored", stage


def cmd_traffic(a):
    f = dispatches(one(a.fetch, "counter_collection.csv"), "FETCH_SIZE")
    w = dispatches(one(a.write, "counter_collection.csv"), "WRITE_SIZE")
    if [d[1] for d in f] != [d[1] for d in w]:
        sys.exit("the FETCH_SIZE and WRITE_SIZE runs dispatched different kernels")
    copies = [i for i, d in enumerate(f) if d[1].startswith("__amd_rocclr_copyBuffer")]
    ic = max(copies, key=lambda i: f[i][2])
    ff = COPY_BYTES / (f[ic][2] * 1024.0)
    fw = COPY_BYTES / (w[ic][2] * 1024.0)
    lines_common = [
        "Calibration: the 4 GiB device-to-device copy (__amd_rocclr_copyBuffer, grid "
        f"{f[ic][3]}) reads and writes 4 GiB; it reports FETCH_SIZE {f[ic][2]:.0f} KiB and "
        f"WRITE_SIZE {w[ic][2]:.0f} KiB.",
        f"Scale factors: bytes read = FETCH_SIZE x 1024 x {ff:.3f}; bytes written = "
        f"WRITE_SIZE x 1024 x {fw:.3f}.", ""]
    for wl, lo, hi in ((a.before, 0, ic), (a.after, ic + 1, len(f))):
        if not wl:
            continue
        out = [f"# HBM traffic of the M2L stage per evaluation pass: {wl} (rocprofv3 --pmc)",
               "", "Commands (MI355X, one GPU, repository root), one counter per run:",
               "  cd /tmp && export TMPDIR=/tmp && cd /root/repo"]
        out += ["  " + c for c in a.cmd]
        out += [f"  python3 profile_summary.py traffic <fetch dir> <write dir> "
                f"--before {a.before} --after {a.after or ''}", ""] + lines_common
        keys = {}
        out.append("pass  prec  rhs  M2L variant  M2L launches  read GB  write GB  "
                   "(whole pass: read GB, write GB)")
        for n, (s0, seg) in enumerate(passes(f[lo:hi])):
            wseg = w[lo + s0: lo + s0 + len(seg)]
            prec, nrhs, var, stage = describe(seg)
            rd = sum(seg[i][2] for i in stage) * 1024 * ff
            wr = sum(wseg[i][2] for i in stage) * 1024 * fw
            prd = sum(d[2] for d in seg) * 1024 * ff
            pwr = sum(d[2] for d in wseg) * 1024 * fw
            nl = sum(1 for i in stage if is_m2l(seg[i][1]))
            out.append(f"{n:4d}  {prec}  {nrhs:4d}  {var:11s}  {nl:12d}  {rd / 1e9:7.3f}  "
                       f"{wr / 1e9:8.3f}  ({prd / 1e9:.3f}, {pwr / 1e9:.3f})")
            if var != "none":
                key = f"{var}_{prec}_x{nrhs}"
                keys[key] = (rd, wr, seg, wseg, stage)
        out.append("")
        for key, (rd, wr, seg, wseg, stage) in keys.items():
            out.append(f"{key} (last pass of this kind), M2L stage by kernel:")
            per = collections.OrderedDict()
            for i in stage:
                p = per.setdefault(seg[i][1], [0, 0.0, 0.0])
                p[0] += 1
                p[1] += seg[i][2] * 1024 * ff
                p[2] += wseg[i][2] * 1024 * fw
            for k, (cnt, r_, w_) in per.items():
                out.append(f"  {k:40s} x{cnt:4d}  read {r_ / 1e9:8.3f} GB  write {w_ / 1e9:8.3f} GB")
        out.append("")
        for key, (rd, wr, *_r) in keys.items():
            out.append(f"{key}_read_bytes_per_pass: {rd:.6e}")
            out.append(f"{key}_write_bytes_per_pass: {wr:.6e}")
            out.append(f"{key}_bytes_per_pass: {rd + wr:.6e}")
        path = os.path.join(ROOT, "profiles", f"m2l_traffic_{wl}.txt")
        open(path, "w").write("\n".join(out) + "\n")
        print(f"wrote {path}")


def cmd_kernels(a):
    st = rows(one(a.dir, "kernel_stats.csv"))
    ds = dispatches(one(a.dir, "kernel_trace.csv"))
    out = [f"# Kernel durations: {a.title} (rocprofv3 --kernel-trace --stats)", "",
           "Command (MI355X, one GPU, repository root):",
           "  cd /tmp && export TMPDIR=/tmp && cd /root/repo"]
    out += ["  " + c for c in a.cmd]
    out += [f"  python3 profile_summary.py kernels <dir> --title '{a.title}'", "",
            "All dispatches of the run (every section of the bench command):", "",
            f"{'kernel':48s} {'calls':>7s} {'total ms':>10s} {'avg us':>10s} {'min us':>9s} "
            f"{'max us':>9s}  share"]
    st.sort(key=lambda r: -float(r["TotalDurationNs"]))
    for r in st[:a.top]:
        out.append(f"{short(r['Name'] + '(') :48s} {int(r['Calls']):7d} "
                   f"{float(r['TotalDurationNs']) / 1e6:10.3f} {float(r['AverageNs']) / 1e3:10.2f} "
                   f"{float(r['MinNs']) / 1e3:9.2f} {float(r['MaxNs']) / 1e3:9.2f}  "
                   f"{float(r['Percentage']):5.2f}%")
    out += ["", "Per evaluation pass: kernel time of the M2L stage (sum of its kernels' "
            "durations) and of the whole pass:", "",
            "pass  prec  rhs  M2L variant  M2L launches  avg M2L launch us  M2L stage ms  "
            "pass kernels ms"]
    for n, (s0, seg) in enumerate(passes(ds)):
        prec, nrhs, var, stage = describe(seg)
        m2l = [seg[i][2] for i in stage if is_m2l(seg[i][1])]
        tot = sum(seg[i][2] for i in stage)
        avg = sum(m2l) / len(m2l) / 1e3 if m2l else 0.0
        out.append(f"{n:4d}  {prec}  {nrhs:4d}  {var:11s}  {len(m2l):12d}  {avg:17.2f}  "
                   f"{tot / 1e6:12.3f}  {sum(d[2] for d in seg) / 1e6:15.3f}")
    open(a.out, "w").write("\n".join(out) + "\n")
    print(f"wrote {a.out}")


def cmd_counters(a):
    rs = rows(one(a.dir, "counter_collection.csv"))
    per = collections.defaultdict(lambda: collections.defaultdict(float))
    calls = collections.defaultdict(set)
    for r in rs:
        k = short(r["Kernel_Name"])
        per[k][r["Counter_Name"]] += float(r["Counter_Value"])
        calls[k].add(r["Dispatch_Id"])
    names = sorted({c for v in per.values() for c in v})
    tanote = " (TA_BUSY_avr: averaged over TA instances, then summed)" if "TA_BUSY_avr" in names else ""
    out = [f"# Instruction counters per kernel: {a.title} (rocprofv3 --pmc)", "",
           "Command (MI355X, one GPU, repository root):",
           "  cd /tmp && export TMPDIR=/tmp && cd /root/repo"]
    out += ["  " + c for c in a.cmd]
    out += [f"  python3 profile_summary.py counters <dir> --title '{a.title}'", "",
            f"Sums over every dispatch of the run{tanote}.  GRBM_GUI_ACTIVE is reported summed over the "
            f"{a.xcds} XCDs, so kernel cycles = GRBM_GUI_ACTIVE / {a.xcds}.  MFMA busy = "
            f"SQ_VALU_MFMA_BUSY_CYCLES / (cycles x {a.simds} SIMDs) (rocprofv3's MfmaUtil); VALU "
            f"busy = 4 x SQ_INSTS_VALU / (cycles x {a.simds} SIMDs) (a wave64 VALU instruction "
            "holds a 16-lane SIMD 4 cycles; an estimate: some instructions take longer).", ""]
    hit = "TCC_HIT_sum" in names and "TCC_MISS_sum" in names
    sq = "SQ_INSTS_VALU" in names
    if hit:
        out += ["L2 hit = TCC_HIT_sum / (TCC_HIT_sum + TCC_MISS_sum); TA busy = TA_BUSY_avr / kernel cycles.", ""]
    head = f"{'kernel':44s} {'calls':>6s} " + " ".join(f"{c:>24s}" for c in names)
    ta = "TA_BUSY_avr" in names
    out.append(head + ("   MFMA busy  VALU busy" if sq else "") + ("  L2 hit" if hit else "")
               + ("  TA busy" if ta else ""))
    order = sorted(per, key=lambda k: -per[k].get("GRBM_GUI_ACTIVE", 0))
    for k in order[:a.top]:
        v = per[k]
        g = v.get("GRBM_GUI_ACTIVE", 0) / a.xcds * a.simds
        mb = 100 * v.get("SQ_VALU_MFMA_BUSY_CYCLES", 0) / g if g else 0
        vi = 400 * v.get("SQ_INSTS_VALU", 0) / g if g else 0
        row = f"{k:44s} {len(calls[k]):6d} " + " ".join(f"{v.get(c, 0):24.0f}" for c in names)
        if sq:
            row += f"   {mb:8.2f}%  {vi:9.2f}%"
        if hit:
            hm = v.get("TCC_HIT_sum", 0) + v.get("TCC_MISS_sum", 0)
            row += f"  {100 * v.get('TCC_HIT_sum', 0) / hm:5.1f}%" if hm else "       -"
        if ta:
            cyc = v.get("GRBM_GUI_ACTIVE", 0) / a.xcds
            row += f"  {100 * v.get('TA_BUSY_avr', 0) / cyc:6.1f}%" if cyc else "        -"
        out.append(row)
    open(a.out, "w").write("\n".join(out) + "\n")
    print(f"wrote {a.out}")


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    sub = ap.add_subparsers(dest="what", required=True)
    t = sub.add_parser("traffic")
    t.add_argument("fetch")
    t.add_argument("write")
    t.add_argument("--before", required=True)
    t.add_argument("--after", default="")
    t.add_argument("--cmd", action="append", default=[])
    k = sub.add_parser("kernels")
    k.add_argument("dir")
    k.add_argument("--out", required=True)
    k.add_argument("--title", required=True)
    k.add_argument("--cmd", action="append", default=[])
    k.add_argument("--top", type=int, default=30)
    c = sub.add_parser("counters")
    c.add_argument("dir")
    c.add_argument("--out", required=True)
    c.add_argument("--title", required=True)
    c.add_argument("--cmd", action="append", default=[])
    c.add_argument("--top", type=int, default=25)
    c.add_argument("--simds", type=int, default=1024, help="SIMDs of the device (256 CUs x 4)")
    c.add_argument("--xcds", type=int, default=8, help="XCDs of the device")
    a = ap.parse_args()
    {"traffic": cmd_traffic, "kernels": cmd_kernels, "counters": cmd_counters}[a.what](a)


if __name__ == "__main__":
    main()
