"""Summaries of rocprofv3 output for the voxelizer workloads.

    python -m voxvizthesis_amd.profiling stats <kt_kernel_stats.csv>
    python -m voxvizthesis_amd.profiling trace <kt_kernel_trace.csv> \
        [--after-last KERNEL] [--per-step KERNEL]
    python -m voxvizthesis_amd.profiling traffic <fetch.csv> <write.csv> \
        --known-fetch BYTES --known-write BYTES [--json out.json]

`traffic` turns the FETCH_SIZE / WRITE_SIZE counters (KiB per dispatch) into
HBM bytes per launch.  With --known-copy BYTES (a run of bench.py
--copy-calibration) each counter is scaled by the ratio it shows on the plain
copy kernel k_copy_u128, which reads and writes exactly BYTES; the ratios the
same copy shows with its 256-byte blocks in a scattered order
(k_copy_scattered) are reported beside them.  Otherwise
the WRITE_SIZE scale comes from the largest k_fill_u32 dispatch, which
writes exactly the S^3 occupancy bitmap (--known-write), and applies to
FETCH_SIZE too unless --known-fetch names the bytes the largest
k_scan_reduce dispatch reads.
"""
from __future__ import annotations

import argparse
import collections
import csv
import json
import re


def kname(s: str) -> str:
    # rocPRIM's sort kernels: the part of the template that names the pass
    m = re.search(r"wrapped_(\w+?)_config", s)
    if s.startswith("void rocprim") and m:
        return "rocprim " + m.group(1)
    for prefix in ("voxmi::(anonymous namespace)::", "(anonymous namespace)::"):
        if s.startswith(prefix):
            s = s[len(prefix):]
    return s.split("(")[0]


def load_counter(path: str, counter: str):
    per = collections.defaultdict(float)
    names = {}
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] != counter:
            continue
        d = int(r["Dispatch_Id"])
        per[d] += float(r["Counter_Value"])
        names[d] = kname(r["Kernel_Name"])
    by_kernel = collections.defaultdict(list)
    for d, v in sorted(per.items()):
        by_kernel[names[d]].append(v * 1024.0)      # KiB -> bytes
    return by_kernel


def stats(path: str, prefix: str = "") -> str:
    """Markdown table of a kernel-stats CSV; prefix keeps only the kernels
    whose short name starts with it (k_srv, k_ray ...)."""
    rows = [r for r in csv.DictReader(open(path)) if kname(r["Name"]).startswith(prefix)]
    out = ["| kernel | calls | avg us | total % |", "|---|---|---|---|"]
    for r in rows:
        out.append(f"| {kname(r['Name'])} | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | "
                   f"{float(r['Percentage']):.1f} |")
    return "\n".join(out)


def traffic(fetch_csv, write_csv, known_fetch, known_write, steps=1, known_copy=0.0):
    f = load_counter(fetch_csv, "FETCH_SIZE")
    w = load_counter(write_csv, "WRITE_SIZE")
    if known_copy:
        fw = known_copy / max(w["k_copy_u128"])
        ff = known_copy / max(f["k_copy_u128"])
    else:
        fw = known_write / max(w["k_fill_u32"])
        ff = known_fetch / max(f["k_scan_reduce"]) if known_fetch else fw
    res = {"fetch_scale": ff, "write_scale": fw, "kernels": {}}
    if known_copy and f.get("k_copy_scattered") and w.get("k_copy_scattered"):
        # the second reference point: the same bytes copied in 256-byte
        # blocks in a scattered order; its scales bound how far the plain
        # copy's scales carry over to kernels that gather
        res["scattered_fetch_scale"] = known_copy / max(f["k_copy_scattered"])
        res["scattered_write_scale"] = known_copy / max(w["k_copy_scattered"])
    for k in sorted(set(f) | set(w)):
        fs = f.get(k, [])
        ws = w.get(k, [])
        tot = sum(fs) * ff + sum(ws) * fw
        n = max(len(fs), len(ws), 1)
        res["kernels"][k] = {"launches": n, "fetch_bytes": sum(fs) * ff,
                             "write_bytes": sum(ws) * fw, "bytes_per_launch": tot / n}
    return res


def trace_stats(path: str, after_last: str = "", per_step: str = "") -> str:
    """Markdown table of a kernel-trace CSV: the dispatches after the last
    dispatch of after_last (a kernel that ends one phase of a run), per
    kernel calls, average and, with per_step (a kernel dispatched once per
    step), the time per step and its share."""
    rows = [(int(r["Start_Timestamp"]), int(r["End_Timestamp"]), kname(r["Kernel_Name"]))
            for r in csv.DictReader(open(path))]
    rows.sort()
    if after_last:
        last = max((i for i, r in enumerate(rows) if r[2] == after_last), default=-1)
        rows = rows[last + 1:]
    tot = collections.defaultdict(float)
    calls = collections.Counter()
    for s0, e0, n in rows:
        tot[n] += (e0 - s0) / 1e3
        calls[n] += 1
    steps = calls[per_step] if per_step else 0
    allus = sum(tot.values())
    out = ["| kernel | calls | avg us | us per step | share % |", "|---|---|---|---|---|"]
    for n in sorted(tot, key=lambda k: -tot[k]):
        ps = f"{tot[n] / steps:.1f}" if steps else ""
        out.append(f"| {n} | {calls[n]} | {tot[n] / calls[n]:.1f} | {ps} | "
                   f"{100.0 * tot[n] / allus:.1f} |")
    if steps:
        out.append(f"| total ({steps} steps) | | | {allus / steps:.1f} | 100.0 |")
    return "\n".join(out)


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("stats")
    a.add_argument("csv")
    a.add_argument("--prefix", default="")
    t = sub.add_parser("trace")
    t.add_argument("csv")
    t.add_argument("--after-last", default="")
    t.add_argument("--per-step", default="")
    b = sub.add_parser("traffic")
    b.add_argument("fetch")
    b.add_argument("write")
    b.add_argument("--known-fetch", type=float, default=0.0)
    b.add_argument("--known-write", type=float, default=0.0)
    b.add_argument("--known-copy", type=float, default=0.0)
    b.add_argument("--json")
    args = ap.parse_args()
    if args.cmd == "stats":
        print(stats(args.csv, args.prefix))
    elif args.cmd == "trace":
        print(trace_stats(args.csv, args.after_last, args.per_step))
    else:
        res = traffic(args.fetch, args.write, args.known_fetch, args.known_write,
                      known_copy=args.known_copy)
        if args.json:
            with open(args.json, "w") as fh:
                json.dump(res, fh, indent=1)
        for k, v in sorted(res["kernels"].items(), key=lambda kv: -kv[1]["bytes_per_launch"]):
            print(f"{k:20s} launches={v['launches']:3d} bytes/launch={v['bytes_per_launch']:.4g}")
        print(f"scales: fetch {res['fetch_scale']:.4f} write {res['write_scale']:.4f}")
        if "scattered_fetch_scale" in res:
            print(f"scattered copy: fetch {res['scattered_fetch_scale']:.4f} "
                  f"write {res['scattered_write_scale']:.4f}")


if __name__ == "__main__":
    main()
