"""Summaries of rocprofv3 CSV output as short markdown (for the profiles folder).

    python -m h8pcg.profsum stats <dir with *_kernel_stats.csv> [--top 12]
    python -m h8pcg.profsum traffic --fetch <dir> --write <dir> [--calib-bytes 268435456]
    python -m h8pcg.profsum launches <dir with *_kernel_trace.csv> --kernel NAME [--last 2]

``stats`` prints the per-kernel table of a ``--kernel-trace --stats`` run.
``traffic`` reads two ``--pmc FETCH_SIZE`` / ``--pmc WRITE_SIZE`` runs of
``python -m h8pcg.kbench``, averages each counter per dispatch (KiB), and scales
each by what the same counter reads for the calibration kernel (torch's
elementwise ``neg`` over --calib-bytes bytes, which reads and writes exactly that
many bytes).  It prints one ``traffic-json:`` line per h8 kernel, in bytes per
launch, that bench.py reads back.  ``launches`` prints the durations of the last
--last dispatches of the kernels whose name contains NAME, in dispatch order
(for a kernel whose launches differ in length, such as a solver that runs a
whole solve phase per launch, the dispatches of one timing run).
"""
from __future__ import annotations

import argparse
import collections
import csv
import glob
import json
import os


def _one(pattern):
    files = sorted(glob.glob(pattern, recursive=True))
    if not files:
        raise SystemExit(f"profsum: nothing matches {pattern}")
    return files[0]


def short_name(name):
    n = name.split("(")[0] if not name.startswith("void ") else name[5:].split("(")[0]
    return n.replace("h8::", "")


def stats(path, top):
    f = _one(os.path.join(path, "**", "*kernel_stats.csv"))
    rows = list(csv.DictReader(open(f)))
    print("| kernel | calls | avg us | min us | max us | % of kernel time |")
    print("|---|---:|---:|---:|---:|---:|")
    for r in rows[:top]:
        print(f"| `{short_name(r['Name'])[:70]}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.2f} | "
              f"{float(r['MinNs']) / 1e3:.2f} | {float(r['MaxNs']) / 1e3:.2f} | {float(r['Percentage']):.1f} |")


def per_dispatch(path, counter):
    f = _one(os.path.join(path, "**", "*counter_collection.csv"))
    acc = collections.defaultdict(lambda: collections.defaultdict(float))
    for r in csv.DictReader(open(f)):
        if r["Counter_Name"] != counter:
            continue
        acc[r["Kernel_Name"]][r["Dispatch_Id"]] += float(r["Counter_Value"])
    return {k: sum(v.values()) / len(v) for k, v in acc.items()}


def traffic(fetch_dir, write_dir, calib_bytes):
    fe = per_dispatch(fetch_dir, "FETCH_SIZE")
    wr = per_dispatch(write_dir, "WRITE_SIZE")
    calib = [k for k in fe if "neg_kernel" in k]
    if not calib:
        raise SystemExit("profsum: the calibration kernel (neg) is missing from the FETCH run")
    ck = calib[0]
    f_scale = calib_bytes / (fe[ck] * 1024.0)
    w_scale = calib_bytes / (wr[ck] * 1024.0)
    print(f"Calibration kernel `neg` (float32, {calib_bytes} B read and written): FETCH_SIZE reads "
          f"{fe[ck]:.0f} KiB, WRITE_SIZE {wr[ck]:.0f} KiB per dispatch -> scale FETCH by {f_scale:.3f}, "
          f"WRITE by {w_scale:.3f}.\n")
    print("| kernel | FETCH_SIZE KiB | WRITE_SIZE KiB | HBM bytes read (scaled) | written (scaled) | total |")
    print("|---|---:|---:|---:|---:|---:|")
    out = []
    for k in fe:
        if "h8::" not in k or k not in wr:
            continue
        rb = fe[k] * 1024.0 * f_scale
        wb = wr[k] * 1024.0 * w_scale
        print(f"| `{short_name(k)}` | {fe[k]:.0f} | {wr[k]:.0f} | {rb:.0f} | {wb:.0f} | {rb + wb:.0f} |")
        out.append({"kernel": short_name(k), "read_bytes": rb, "write_bytes": wb, "bytes_per_launch": rb + wb})
    print()
    for o in out:
        print("traffic-json: " + json.dumps(o))


def launches(path, kernel, last):
    f = _one(os.path.join(path, "**", "*kernel_trace.csv"))
    rows = [r for r in csv.DictReader(open(f)) if kernel in r["Kernel_Name"]]
    if not rows:
        raise SystemExit(f"profsum: no dispatch of a kernel named like {kernel!r}")
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    print(f"{len(rows)} dispatches of `{short_name(rows[0]['Kernel_Name'])}`; the last {min(last, len(rows))}:\n")
    print("| dispatch | duration us |")
    print("|---:|---:|")
    for r in rows[-last:]:
        print(f"| {r['Dispatch_Id']} | {(int(r['End_Timestamp']) - int(r['Start_Timestamp'])) / 1e3:.2f} |")


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("dir")
    s.add_argument("--top", type=int, default=12)
    t = sub.add_parser("traffic")
    t.add_argument("--fetch", required=True)
    t.add_argument("--write", required=True)
    t.add_argument("--calib-bytes", type=int, default=256 << 20)
    ln = sub.add_parser("launches")
    ln.add_argument("dir")
    ln.add_argument("--kernel", required=True)
    ln.add_argument("--last", type=int, default=2)
    a = ap.parse_args(argv)
    if a.cmd == "stats":
        stats(a.dir, a.top)
    elif a.cmd == "launches":
        launches(a.dir, a.kernel, a.last)
    else:
        traffic(a.fetch, a.write, a.calib_bytes)


if __name__ == "__main__":
    main()
