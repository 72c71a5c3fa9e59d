"""Summaries of rocprofv3 CSV output for the PBF substep kernels.

  python rocprof_summary.py kernels <trace_dir> --cmd "<command>"
      per-kernel calls / average / share from <trace_dir>/run_kernel_stats.csv
  python rocprof_summary.py counters <pmc_dir> --cmd "<command>"
      per-kernel average of every counter of one --pmc run (+ VALU per wave)
  python rocprof_summary.py derived <pmcA_dir> <pmcB_dir>
      VALU instructions, active lanes, VALU-busy share, loads and LDS
      instructions per wave of the gather kernels from the two SQ passes
  python rocprof_summary.py caches --run LABEL=<cache_dir>,<fetch_dir> [--run ...] --fetch-factor F
                            --cmd "<command>" [--cmd ...]
      per-kernel L1 and L2 hit rates (TCP_TOTAL_CACHE_ACCESSES_sum,
      TCP_TCC_READ_REQ_sum, TCC_HIT_sum, TCC_MISS_sum) and HBM read bytes
      (FETCH_SIZE x F) of several labelled runs side by side
  python rocprof_summary.py traffic <fetch_dir> <write_dir> --calib-fetch <dir> --calib-write <dir>
                            --calib-bytes B --cmd "<command>" [--json out.json]
      HBM bytes per dispatch of each kernel: FETCH_SIZE and WRITE_SIZE (KiB per
      dispatch) scaled by the factor that maps the same counters of a plain copy
      kernel of B bytes read and B bytes written onto B.

Kernels are grouped into the passes of a substep by name.
"""
import argparse
import collections
import csv
import json
import os
import re
import sys

PASS_OF = [
    ("k_predict", "predict"), ("k_radix", "sort"), ("k_scan", "sort"), ("k_iota", "sort"),
    ("k_clear_cells", "tables"), ("k_build_cells", "tables"), ("k_reorder", "tables"), ("k_order", "tables"),
    ("k_bin", "tables"), ("k_fill", "tables"), ("k_lod", "lod"), ("k_lambda", "lambda"), ("k_deltap", "deltap"),
    ("k_apply_sdf", "deltap"), ("k_integrate", "integrate"), ("k_vorticity", "vorticity"), ("k_contact", "contacts"),
    ("k_seg_", "contacts"), ("k_vertex_sum", "contacts"), ("k_sdf", "mesh"), ("k_mark", "mesh"), ("k_refit", "mesh"),
    ("k_update_tris", "mesh"), ("k_dirty_init", "mesh"), ("k_edges", "adjoint lists"), ("k_transpose", "adjoint lists"),
    ("k_tstats", "adjoint lists"), ("k_zero_tcnt", "adjoint lists"), ("k_rec_", "adjoint replay"),
    ("k_predict64", "adjoint replay"), ("k_reorder64", "adjoint replay"), ("k_rev_", "adjoint reverse"),
    ("k_absmax", "adjoint reverse"), ("k_compare", "adjoint replay"), ("k_reduce2", "adjoint reverse"),
]


def short_name(name: str) -> str:
    m = re.search(r"(k_[A-Za-z0-9_]+)(<[^()]*>)?\(", name)
    if m:
        return m.group(1) + (m.group(2) or "")
    return name.split("(")[0][:60]


def pass_of(short: str) -> str:
    for prefix, p in PASS_OF:
        if short.startswith(prefix):
            return p
    return "other"


def kernels(args):
    rows = list(csv.DictReader(open(os.path.join(args.dir, "run_kernel_stats.csv"))))
    total = sum(float(r["TotalDurationNs"]) for r in rows)
    print(f"# rocprofv3 kernel statistics\n\nCommand:\n\n    {args.cmd}\n")
    print("| kernel | pass | calls | avg us | min us | max us | share |")
    print("|---|---|---|---|---|---|---|")
    by_pass = collections.defaultdict(float)
    for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"])):
        s = short_name(r["Name"])
        p = pass_of(s)
        by_pass[p] += float(r["TotalDurationNs"])
        print(f"| `{s}` | {p} | {r['Calls']} | {float(r['AverageNs']) / 1e3:.2f} | {float(r['MinNs']) / 1e3:.2f} | "
              f"{float(r['MaxNs']) / 1e3:.2f} | {float(r['TotalDurationNs']) / total:.4f} |")
    print("\n| pass | kernel time share |\n|---|---|")
    for p, t in sorted(by_pass.items(), key=lambda kv: -kv[1]):
        print(f"| {p} | {t / total:.4f} |")


def counter_means(d: str, counter: str):
    acc = collections.defaultdict(list)
    for r in csv.DictReader(open(os.path.join(d, "run_counter_collection.csv"))):
        if r["Counter_Name"] == counter:
            acc[short_name(r["Kernel_Name"])].append(float(r["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in acc.items()}, {k: len(v) for k, v in acc.items()}


def traffic(args):
    fetch, calls = counter_means(args.fetch_dir, "FETCH_SIZE")
    write, _ = counter_means(args.write_dir, "WRITE_SIZE")
    cf, _ = counter_means(args.calib_fetch, "FETCH_SIZE")
    cw, _ = counter_means(args.calib_write, "WRITE_SIZE")
    copy = [k for k in cf if re.search(args.calib_kernel, k)]
    if not copy:
        sys.exit(f"no calibration kernel matching {args.calib_kernel!r} in {sorted(cf)}")
    ck = max(copy, key=lambda k: cf[k])
    kr = args.calib_bytes / (cf[ck] * 1024.0)
    kw = args.calib_bytes / (cw[ck] * 1024.0)
    print(f"# HBM traffic per dispatch (FETCH_SIZE / WRITE_SIZE)\n\nCommand:\n\n    {args.cmd}\n")
    print(f"Calibration: `{ck}` copying {args.calib_bytes} bytes read FETCH_SIZE {cf[ck]:.1f} KiB and WRITE_SIZE "
          f"{cw[ck]:.1f} KiB per dispatch, so bytes = FETCH_SIZE x 1024 x {kr:.4f} and WRITE_SIZE x 1024 x {kw:.4f}.\n")
    print("| kernel | pass | dispatches | FETCH KiB | WRITE KiB | read MB | written MB | total MB |")
    print("|---|---|---|---|---|---|---|---|")
    out = {"calibration": {"kernel": ck, "bytes": args.calib_bytes, "read_factor": kr, "write_factor": kw},
           "kernels": {}}
    for k in sorted(fetch, key=lambda k: -fetch[k]):
        rd = fetch[k] * 1024.0 * kr
        wr = write.get(k, 0.0) * 1024.0 * kw
        out["kernels"][k] = {"pass": pass_of(k), "read_bytes": rd, "write_bytes": wr, "bytes_per_launch": rd + wr}
        print(f"| `{k}` | {pass_of(k)} | {calls[k]} | {fetch[k]:.1f} | {write.get(k, 0.0):.1f} | {rd / 1e6:.2f} | "
              f"{wr / 1e6:.2f} | {(rd + wr) / 1e6:.2f} |")
    if args.json:
        with open(args.json, "w") as f:
            json.dump(out, f, indent=1)


def counters(args):
    """Per-kernel averages of every counter of one --pmc run, and VALU
    instructions per wave when SQ_WAVES and SQ_INSTS_VALU are among them."""
    acc = collections.defaultdict(lambda: collections.defaultdict(list))
    for r in csv.DictReader(open(os.path.join(args.dir, "run_counter_collection.csv"))):
        acc[short_name(r["Kernel_Name"])][r["Counter_Name"]].append(float(r["Counter_Value"]))
    names = sorted({c for k in acc.values() for c in k})
    print(f"# Hardware counters per dispatch\n\nCommand:\n\n    {args.cmd}\n")
    extra = "SQ_WAVES" in names and "SQ_INSTS_VALU" in names
    print("| kernel | dispatches | " + " | ".join(names) + (" | VALU inst / wave |" if extra else " |"))
    print("|---|---|" + "---|" * len(names) + ("---|" if extra else ""))
    rows = sorted(acc.items(), key=lambda kv: -sum(kv[1].get("SQ_INSTS_VALU", [0.0])))
    for k, cs in rows:
        means = {c: sum(v) / len(v) for c, v in cs.items()}
        line = f"| `{k}` | {len(next(iter(cs.values())))} | " + " | ".join(f"{means.get(c, 0.0):.4g}" for c in names)
        if extra:
            w = means.get("SQ_WAVES", 0.0)
            line += f" | {means.get('SQ_INSTS_VALU', 0.0) / w:.0f}" if w else " | -"
        print(line + " |")


def derived(args):
    """Per-wave rates of the gather kernels from two --pmc runs (A: SQ_WAVES,
    SQ_INSTS_VALU, SQ_ACTIVE_INST_VALU, SQ_BUSY_CYCLES, SQ_INSTS_VMEM_RD,
    SQ_INSTS_LDS; B: SQ_THREAD_CYCLES_VALU), on a device of `--simds` SIMDs
    and `--engines` shader engines."""
    a = collections.defaultdict(lambda: collections.defaultdict(list))
    for d in (args.dir_a, args.dir_b):
        for r in csv.DictReader(open(os.path.join(d, "run_counter_collection.csv"))):
            a[short_name(r["Kernel_Name"])][r["Counter_Name"]].append(float(r["Counter_Value"]))
    print("| kernel | VALU inst / wave | thread-cycles / VALU inst | VALU-active wave cycles / SIMD cycles | "
          "vector loads / wave | LDS inst / wave |")
    print("|---|---|---|---|---|---|")
    m = {k: {c: sum(v) / len(v) for c, v in cs.items()} for k, cs in a.items()}
    for k in sorted(m, key=lambda k: -m[k].get("SQ_INSTS_VALU", 0.0)):
        c = m[k]
        w, v = c.get("SQ_WAVES", 0.0), c.get("SQ_INSTS_VALU", 0.0)
        if not w or not v or not re.match(args.match, k):
            continue
        lanes = c.get("SQ_THREAD_CYCLES_VALU", 0.0) / v
        cycles = c.get("SQ_BUSY_CYCLES", 0.0) / args.engines
        busy = 4.0 * c.get("SQ_ACTIVE_INST_VALU", 0.0) / (cycles * args.simds) if cycles else 0.0
        print(f"| `{k}` | {v / w:.0f} | {lanes:.1f} of 64 ({lanes / 64:.0%}) | {busy:.0%} | "
              f"{c.get('SQ_INSTS_VMEM_RD', 0.0) / w:.0f} | {c.get('SQ_INSTS_LDS', 0.0) / w:.0f} |")


def caches(args):
    """L1 / L2 hit rates and HBM reads of each kernel across labelled runs."""
    runs = []
    for spec in args.run:
        label, dirs = spec.split("=", 1)
        cdir, fdir = dirs.split(",")
        hit, calls = counter_means(cdir, "TCC_HIT_sum")
        miss, _ = counter_means(cdir, "TCC_MISS_sum")
        l1, _ = counter_means(cdir, "TCP_TOTAL_CACHE_ACCESSES_sum")
        l1m, _ = counter_means(cdir, "TCP_TCC_READ_REQ_sum")
        fetch, _ = counter_means(fdir, "FETCH_SIZE")
        runs.append((label, hit, miss, l1, l1m, fetch, calls))
    print("# L1 / L2 hit rates and HBM reads per dispatch\n\nCommands:\n")
    for c in args.cmd:
        print(f"    {c}")
    print(f"\nHBM read bytes = FETCH_SIZE (KiB) x {args.fetch_factor:.1f}, the factor of the copy-kernel calibration.\n")
    print("| kernel | pass | run | dispatches | L1 hit % | L2 hit % | L2 requests | HBM read MB |")
    print("|---|---|---|---|---|---|---|---|")
    names = sorted({k for r in runs for k in r[1]}, key=lambda k: (-max(r[5].get(k, 0.0) for r in runs), k))
    for k in names:
        for label, hit, miss, l1, l1m, fetch, calls in runs:
            if k not in hit:
                continue
            req = hit[k] + miss.get(k, 0.0)
            l2 = 100.0 * hit[k] / req if req else 0.0
            l1r = 100.0 * (1.0 - l1m.get(k, 0.0) / l1[k]) if l1.get(k) else 0.0
            mb = fetch.get(k, 0.0) * args.fetch_factor / 1e6
            print(f"| `{k}` | {pass_of(k)} | {label} | {calls[k]} | {l1r:.1f} | {l2:.1f} | {req:.4g} | {mb:.2f} |")


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="what", required=True)
    k = sub.add_parser("kernels")
    k.add_argument("dir")
    k.add_argument("--cmd", required=True)
    t = sub.add_parser("traffic")
    t.add_argument("fetch_dir")
    t.add_argument("write_dir")
    t.add_argument("--calib-fetch", required=True)
    t.add_argument("--calib-write", required=True)
    t.add_argument("--calib-bytes", type=float, required=True)
    t.add_argument("--calib-kernel", default="elementwise|Functor|copy")
    t.add_argument("--cmd", required=True)
    t.add_argument("--json")
    c = sub.add_parser("counters")
    c.add_argument("dir")
    c.add_argument("--cmd", required=True)
    dv = sub.add_parser("derived")
    dv.add_argument("dir_a")
    dv.add_argument("dir_b")
    dv.add_argument("--simds", type=int, default=1024)
    dv.add_argument("--engines", type=int, default=32)
    dv.add_argument("--match", default="k_(lambda|deltap|integrate|vorticity|apply_sdf)")
    h = sub.add_parser("caches")
    h.add_argument("--run", action="append", required=True)
    h.add_argument("--fetch-factor", type=float, required=True)
    h.add_argument("--cmd", action="append", required=True)
    a = ap.parse_args()
    {"kernels": kernels, "traffic": traffic, "counters": counters, "caches": caches,
     "derived": derived}[a.what](a)


if __name__ == "__main__":
    main()
