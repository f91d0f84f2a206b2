"""Markdown summaries of rocprofv3 CSV output, for the profiles/ folder.

    python -m gdmr.rocprof_summary stats DIR [--top N] [--title T] [--command C]
        the kernel-trace statistics (<name>_kernel_stats.csv under DIR): calls, average and
        total duration per kernel.
    python -m gdmr.rocprof_summary pmc FETCH_DIR WRITE_DIR [--calib KERNEL] [--top N] ...
        FETCH_SIZE and WRITE_SIZE per dispatch (<name>_counter_collection.csv of two runs, one
        counter each), in KiB, averaged over each kernel's dispatches.  The fetch factor is
        calibrated on KERNEL, which reads exactly the bytes it writes (WRITE_SIZE counts written
        bytes one for one): factor = mean WRITE / mean FETCH of that kernel.
    python -m gdmr.rocprof_summary counters DIR [DIR ...] --counters C1,C2,... [--kernels K,...]
        the mean per dispatch of each listed counter (looked up in the counter_collection.csv of
        every DIR, so counters collected in separate runs combine), per kernel, for the kernels
        listed (default: every kernel that has all the counters).
    --record KERNEL --bytes B --n N --s S --json PATH (pmc only) also writes the calibrated
        traffic per dispatch of KERNEL beside its algorithmic bytes per launch B (from the bench
        line of the same command) into the JSON file the benchmark reads.

Kernel names are shortened to the template-id (namespaces and argument lists dropped).
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import re
import sys


def short_name(name: str) -> str:
    n = re.sub(r"\(anonymous namespace\)::", "", name.strip())
    if n.startswith("void "):
        n = n[5:]
    # Drop the argument list: the last top-level parenthesised group.
    depth, cut = 0, None
    for i, ch in enumerate(n):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            cut = i
            break
    if cut is not None:
        n = n[:cut]
    head, sep, tail = n.partition("<")
    head = head.split("::")[-1]
    tail = re.sub(r"\b(?:dmr::)?(?:blk::)?", "", tail)
    n = head + sep + tail
    return re.sub(r"\s+", " ", n).strip()


def _find(root: str, suffix: str) -> str:
    hits = sorted(glob.glob(os.path.join(root, "**", "*" + suffix), recursive=True))
    if not hits:
        sys.exit(f"no *{suffix} under {root}")
    return hits[0]


def _col(row: dict, *names):
    for nm in names:
        if nm in row:
            return row[nm]
    raise KeyError(names[0])


def kernel_stats(root: str):
    path = _find(root, "kernel_stats.csv")
    out = []
    with open(path) as f:
        for row in csv.DictReader(f):
            out.append(dict(name=short_name(_col(row, "Name", "KernelName")),
                            calls=int(_col(row, "Calls")),
                            total_ns=float(_col(row, "TotalDurationNs")),
                            avg_ns=float(_col(row, "AverageNs")),
                            pct=float(_col(row, "Percentage"))))
    # Launch shapes that differ only in arguments merge under one short name.
    merged = {}
    for r in out:
        m = merged.setdefault(r["name"], dict(name=r["name"], calls=0, total_ns=0.0, pct=0.0))
        m["calls"] += r["calls"]
        m["total_ns"] += r["total_ns"]
        m["pct"] += r["pct"]
    rows = sorted(merged.values(), key=lambda r: -r["total_ns"])
    for r in rows:
        r["avg_ns"] = r["total_ns"] / max(r["calls"], 1)
    return rows, path


def counter_means(root: str, counter: str):
    """{kernel: (dispatches, mean counter value per dispatch)}."""
    path = _find(root, "counter_collection.csv")
    acc = {}
    with open(path) as f:
        for row in csv.DictReader(f):
            if _col(row, "Counter_Name", "Counter-Name") != counter:
                continue
            name = short_name(_col(row, "Kernel_Name", "Kernel-Name"))
            disp = _col(row, "Dispatch_Id", "Dispatch-Id")
            val = float(_col(row, "Counter_Value", "Counter-Value"))
            d = acc.setdefault(name, {})
            d[disp] = d.get(disp, 0.0) + val  # summed over the counter's instances
    return {k: (len(v), sum(v.values()) / len(v)) for k, v in acc.items()}, path


def cmd_stats(a):
    rows, path = kernel_stats(a.dir)
    total = sum(r["total_ns"] for r in rows)
    print(f"# {a.title}\n")
    if a.command:
        print("Command (from the repository root, on one MI355X):\n\n```\n" + a.command +
              "\n```\n")
    print(f"Source: `{os.path.basename(path)}`; {len(rows)} kernels, "
          f"{total / 1e6:.1f} ms of kernel time.\n")
    print("| kernel | calls | avg us | total ms | % |\n|---|---:|---:|---:|---:|")
    for r in rows[:a.top]:
        print(f"| `{r['name']}` | {r['calls']} | {r['avg_ns'] / 1e3:.1f} | "
              f"{r['total_ns'] / 1e6:.1f} | {100.0 * r['total_ns'] / total:.1f} |")


def cmd_pmc(a):
    fetch, fpath = counter_means(a.fetch_dir, "FETCH_SIZE")
    write, wpath = counter_means(a.write_dir, "WRITE_SIZE")
    if a.calib not in fetch or a.calib not in write:
        sys.exit(f"calibration kernel {a.calib} not in both runs")
    factor = write[a.calib][1] / fetch[a.calib][1]
    rows = []
    for name in set(fetch) & set(write):
        nf, f = fetch[name]
        nw, w = write[name]
        rows.append(dict(name=name, disp=nf, fetch=f, fcal=f * factor, write=w,
                         mib=(f * factor + w) / 1024.0))
    rows.sort(key=lambda r: -r["mib"] * r["disp"])
    print(f"# {a.title}\n")
    if a.command:
        print("Commands (from the repository root, one counter per run):\n\n```\n" + a.command +
              "\n```\n")
    print(f"Calibration: `{a.calib}` reads exactly the bytes it writes; its mean WRITE_SIZE is "
          f"{write[a.calib][1]:.0f} KiB and FETCH_SIZE {fetch[a.calib][1]:.0f} KiB per dispatch, "
          f"so the fetch factor is x{factor:.4f} (WRITE_SIZE taken one for one).\n")
    print("Calibrated traffic = FETCH x factor + WRITE, KiB per dispatch, mean over the kernel's "
          "dispatches in the run.\n")
    print("| kernel | dispatches | fetch KiB (raw) | fetch KiB (calibrated) | write KiB | "
          "traffic MiB |\n|---|---:|---:|---:|---:|---:|")
    for r in rows[:a.top]:
        print(f"| `{r['name']}` | {r['disp']} | {r['fetch']:.0f} | {r['fcal']:.0f} | "
              f"{r['write']:.0f} | {r['mib']:.1f} |")
    if a.record:
        hits = [r for r in rows if r["name"] == a.record]
        if not hits:
            sys.exit(f"{a.record} not measured")
        r = hits[0]
        traffic = (r["fcal"] + r["write"]) * 1024.0
        rec = {}
        if os.path.exists(a.json):
            rec = json.load(open(a.json))
        rec[a.record] = {"n": a.n, "s": a.s, "bytes_per_launch": a.bytes,
                         "traffic_bytes_per_launch": traffic, "fetch_factor": factor,
                         "dispatches": r["disp"], "source": a.source}
        with open(a.json, "w") as f:
            json.dump(rec, f, indent=1)
        print(f"\n`{a.record}`: {traffic / 1e6:.1f} MB per dispatch against {a.bytes / 1e6:.1f} MB "
              f"algorithmic per launch ({traffic / a.bytes:.3f}x).")


def cmd_counters(a):
    names = [c for c in a.counters.split(",") if c]
    per = {}
    srcs = []
    for c in names:
        found = None
        for d in a.dirs:
            try:
                means, path = counter_means(d, c)
            except SystemExit:
                continue
            if means:
                found = means
                srcs.append(f"`{c}`: {os.path.relpath(path)}")
                break
        if found is None:
            sys.exit(f"counter {c} not found under {' '.join(a.dirs)}")
        per[c] = found
    kernels = [k for k in (a.kernels.split(";") if a.kernels else []) if k] or \
        sorted(set.intersection(*(set(v) for v in per.values())))
    print(f"# {a.title}\n")
    if a.command:
        print("Commands (from the repository root, one run per group of counters):\n\n```\n" +
              a.command + "\n```\n")
    print("Sources: " + "; ".join(srcs) + ".\n")
    print("Mean per dispatch over the kernel's dispatches in the run.\n")
    print("| kernel | dispatches | " + " | ".join(f"`{c}`" for c in names) + " |")
    print("|---|---:|" + "---:|" * len(names))
    for k in kernels:
        if not all(k in per[c] for c in names):
            continue
        disp = per[names[0]][k][0]
        print(f"| `{k}` | {disp} | " + " | ".join(f"{per[c][k][1]:.4g}" for c in names) + " |")


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    sub = ap.add_subparsers(dest="mode", required=True)
    s = sub.add_parser("stats")
    s.add_argument("dir")
    p = sub.add_parser("pmc")
    p.add_argument("fetch_dir")
    p.add_argument("write_dir")
    p.add_argument("--calib", default="k_scale_cols<float>")
    p.add_argument("--record")
    p.add_argument("--bytes", type=float, default=0.0)
    p.add_argument("--n", type=int, default=0)
    p.add_argument("--s", type=int, default=0)
    p.add_argument("--json", default="profiles/roofline_traffic.json")
    p.add_argument("--source", default="")
    c = sub.add_parser("counters")
    c.add_argument("dirs", nargs="+")
    c.add_argument("--counters", required=True)
    c.add_argument("--kernels", default="", help="short kernel names separated by ';'")
    for q in (s, p, c):
        q.add_argument("--top", type=int, default=40)
        q.add_argument("--title", default="rocprofv3 summary")
        q.add_argument("--command", default="")
    a = ap.parse_args(argv)
    {"stats": cmd_stats, "pmc": cmd_pmc, "counters": cmd_counters}[a.mode](a)


if __name__ == "__main__":
    main()
