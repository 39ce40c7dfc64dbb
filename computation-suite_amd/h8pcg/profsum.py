"""Summaries of rocprofv3 CSV output as short markdown (for the profiles folder).

    python -m h8pcg.profsum stats <dir with *_kernel_stats.csv> [--top 12] [--by-grid]
    python -m h8pcg.profsum traffic --fetch <dir> --write <dir> [--calib-bytes 268435456] [--by-grid]
    python -m h8pcg.profsum launches <dir with *_kernel_trace.csv> --kernel NAME [--last 2]
    python -m h8pcg.profsum runs <dir with *_kernel_trace.csv> --kernel NAME [--grid G] [--min-run 20]
    python -m h8pcg.profsum window <dir with *_kernel_trace.csv> [--marker k_mg_mc] [--top 40]

``stats`` prints the per-kernel table of a ``--kernel-trace --stats`` run.
With ``--by-grid`` both ``stats`` and ``traffic`` keep one row per (kernel,
grid size) -- the kernels of a multigrid cycle run once per level, and the grid
size tells the levels apart.
``traffic`` reads two ``--pmc FETCH_SIZE`` / ``--pmc WRITE_SIZE`` runs of
``python -m h8pcg.kbench``, averages each counter per dispatch (KiB), and scales
each by what the same counter reads for the calibration kernel (torch's
elementwise ``neg`` over --calib-bytes bytes, which reads and writes exactly that
many bytes).  It prints one ``traffic-json:`` line per h8 kernel, in bytes per
launch, that bench.py reads back.  ``launches`` prints the durations of the last
--last dispatches of the kernels whose name contains NAME, in dispatch order
(for a kernel whose launches differ in length, such as a solver that runs a
whole solve phase per launch, the dispatches of one timing run).  ``runs`` finds
the runs of at least --min-run consecutive dispatches of one kernel (name
containing NAME, grid size G if given) -- the back-to-back launches of a timing
graph -- and prints each run's mean duration, for comparison with the HIP-event
average a timing helper reports for the same launches.  ``window`` summarises
the dispatches from the last one whose name contains --marker (k_mg_mc: the
first kernel of a multigrid set-up, so the last solve of a run) to the end of
the trace: wall time, the time kernels ran, and one row per (kernel, grid).
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


def stats(path, top, by_grid=False):
    if by_grid:
        f = _one(os.path.join(path, "**", "*kernel_trace.csv"))
        acc = collections.defaultdict(list)
        for r in csv.DictReader(open(f)):
            acc[(r["Kernel_Name"], int(r["Grid_Size_X"]))].append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
        total = sum(sum(v) for v in acc.values())
        print("| kernel | grid | calls | avg us | min us | max us | % of kernel time |")
        print("|---|---:|---:|---:|---:|---:|---:|")
        for (k, g), v in sorted(acc.items(), key=lambda kv: -sum(kv[1]))[:top]:
            print(f"| `{short_name(k)[:70]}` | {g} | {len(v)} | {sum(v) / len(v) / 1e3:.2f} | {min(v) / 1e3:.2f} | "
                  f"{max(v) / 1e3:.2f} | {100.0 * sum(v) / total:.1f} |")
        return
    f = _one(os.path.join(path, "**", "*kernel_stats.csv"))
    rows = list(csv.DictReader(open(f)))
    print("| kernel | calls | avg us | min us | max us | % of kernel time |")
    print("|---|---:|---:|---:|---:|---:|")
    for r in rows[:top]:
        print(f"| `{short_name(r['Name'])[:70]}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.2f} | "
              f"{float(r['MinNs']) / 1e3:.2f} | {float(r['MaxNs']) / 1e3:.2f} | {float(r['Percentage']):.1f} |")


def per_dispatch(path, counter, by_grid=False):
    f = _one(os.path.join(path, "**", "*counter_collection.csv"))
    acc = collections.defaultdict(lambda: collections.defaultdict(float))
    for r in csv.DictReader(open(f)):
        if r["Counter_Name"] != counter:
            continue
        key = (r["Kernel_Name"], int(r["Grid_Size"])) if by_grid else (r["Kernel_Name"], None)
        acc[key][r["Dispatch_Id"]] += float(r["Counter_Value"])
    return {k: sum(v.values()) / len(v) for k, v in acc.items()}


def traffic(fetch_dir, write_dir, calib_bytes, by_grid=False):
    fe = per_dispatch(fetch_dir, "FETCH_SIZE", by_grid)
    wr = per_dispatch(write_dir, "WRITE_SIZE", by_grid)
    calib = [k for k in fe if "neg_kernel" in k[0]]
    wcal = [k for k in wr if "neg_kernel" in k[0]]
    if not calib or not wcal:
        raise SystemExit("profsum: the calibration kernel (neg) is missing from the FETCH or WRITE run")
    ck, wk = max(calib, key=lambda k: fe[k]), max(wcal, key=lambda k: wr[k])
    f_scale = calib_bytes / (fe[ck] * 1024.0)
    w_scale = calib_bytes / (wr[wk] * 1024.0)
    print(f"Calibration kernel `neg` (float32, {calib_bytes} B read and written): FETCH_SIZE reads "
          f"{fe[ck]:.0f} KiB, WRITE_SIZE {wr[wk]:.0f} KiB per dispatch -> scale FETCH by {f_scale:.3f}, "
          f"WRITE by {w_scale:.3f}.\n")
    grid = "grid | " if by_grid else ""
    print(f"| kernel | {grid}FETCH_SIZE KiB | WRITE_SIZE KiB | HBM bytes read (scaled) | written (scaled) | total |")
    print("|---|" + ("---:|" if by_grid else "") + "---:|---:|---:|---:|---:|")
    out = []
    for k in fe:
        if "h8::" not in k[0] or k not in wr:
            continue
        rb = fe[k] * 1024.0 * f_scale
        wb = wr[k] * 1024.0 * w_scale
        g = f"{k[1]} | " if by_grid else ""
        print(f"| `{short_name(k[0])}` | {g}{fe[k]:.0f} | {wr[k]:.0f} | {rb:.0f} | {wb:.0f} | {rb + wb:.0f} |")
        o = {"kernel": short_name(k[0]), "read_bytes": rb, "write_bytes": wb, "bytes_per_launch": rb + wb}
        if by_grid:
            o["grid"] = k[1]
        out.append(o)
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


def runs(path, kernel, grid, min_run):
    f = _one(os.path.join(path, "**", "*kernel_trace.csv"))
    rows = sorted(csv.DictReader(open(f)), key=lambda r: int(r["Start_Timestamp"]))
    out, cur = [], []
    for r in rows:
        hit = kernel in r["Kernel_Name"] and (grid is None or int(r["Grid_Size_X"]) == grid)
        if hit and cur and r["Kernel_Name"] == cur[-1]["Kernel_Name"] and r["Grid_Size_X"] == cur[-1]["Grid_Size_X"]:
            cur.append(r)
            continue
        if len(cur) >= min_run:
            out.append(cur)
        cur = [r] if hit else []
    if len(cur) >= min_run:
        out.append(cur)
    if not out:
        raise SystemExit(f"profsum: no run of {min_run} consecutive dispatches of {kernel!r}")
    print("| kernel | grid | first dispatch | launches | mean us | min us | max us |")
    print("|---|---:|---:|---:|---:|---:|---:|")
    for run in out:
        d = [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3 for r in run]
        print(f"| `{short_name(run[0]['Kernel_Name'])[:60]}` | {run[0]['Grid_Size_X']} | {run[0]['Dispatch_Id']} | "
              f"{len(d)} | {sum(d) / len(d):.2f} | {min(d):.2f} | {max(d):.2f} |")


def window(path, marker, top):
    f = _one(os.path.join(path, "**", "*kernel_trace.csv"))
    rows = sorted(csv.DictReader(open(f)), key=lambda r: int(r["Start_Timestamp"]))
    idx = [i for i, r in enumerate(rows) if marker in r["Kernel_Name"]]
    if not idx:
        raise SystemExit(f"profsum: no dispatch of {marker}")
    w = rows[idx[-1]:]
    t0 = int(w[0]["Start_Timestamp"])
    t1 = max(int(r["End_Timestamp"]) for r in w)
    acc = collections.defaultdict(list)
    for r in w:
        acc[(r["Kernel_Name"], int(r["Grid_Size_X"]))].append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
    busy = sum(sum(v) for v in acc.values())
    print(f"{len(w)} dispatches from the last {marker}: wall {(t1 - t0) / 1e3:.1f} us, kernels running "
          f"{busy / 1e3:.1f} us ({100.0 * busy / (t1 - t0):.0f}%)")
    print()
    print("| kernel | grid | calls | total us | avg us | % of kernel time |")
    print("|---|---:|---:|---:|---:|---:|")
    for (k, g), v in sorted(acc.items(), key=lambda kv: -sum(kv[1]))[:top]:
        print(f"| `{short_name(k)[:70]}` | {g} | {len(v)} | {sum(v) / 1e3:.1f} | {sum(v) / len(v) / 1e3:.2f} | "
              f"{100.0 * sum(v) / busy:.1f} |")


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("dir")
    s.add_argument("--top", type=int, default=12)
    s.add_argument("--by-grid", action="store_true")
    t = sub.add_parser("traffic")
    t.add_argument("--fetch", required=True)
    t.add_argument("--write", required=True)
    t.add_argument("--calib-bytes", type=int, default=256 << 20)
    t.add_argument("--by-grid", action="store_true")
    ln = sub.add_parser("launches")
    ln.add_argument("dir")
    ln.add_argument("--kernel", required=True)
    ln.add_argument("--last", type=int, default=2)
    rn = sub.add_parser("runs")
    rn.add_argument("dir")
    rn.add_argument("--kernel", required=True)
    rn.add_argument("--grid", type=int, default=None)
    rn.add_argument("--min-run", type=int, default=20)
    wn = sub.add_parser("window")
    wn.add_argument("dir")
    wn.add_argument("--marker", default="k_mg_mc")
    wn.add_argument("--top", type=int, default=40)
    a = ap.parse_args(argv)
    if a.cmd == "stats":
        stats(a.dir, a.top, a.by_grid)
    elif a.cmd == "launches":
        launches(a.dir, a.kernel, a.last)
    elif a.cmd == "runs":
        runs(a.dir, a.kernel, a.grid, a.min_run)
    elif a.cmd == "window":
        window(a.dir, a.marker, a.top)
    else:
        traffic(a.fetch, a.write, a.calib_bytes, a.by_grid)


if __name__ == "__main__":
    main()
