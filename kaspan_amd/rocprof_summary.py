"""Summaries of rocprofv3 output for profiles/ (text only; the raw output stays where rocprofv3 wrote it).

    python -m kaspan_amd.rocprof_summary stats DIR --title T --command C > profiles/NAME_kernel_stats.md
    python -m kaspan_amd.rocprof_summary pmc --fetch DIR --write DIR --copy-fetch DIR --copy-write DIR \
        --copy-bytes B --workload NAME --title T --command C [--command C ...] --json profiles/NAME_pmc_traffic.json \
        > profiles/NAME_pmc_traffic.md

`stats` reads the *_kernel_stats.csv of a `rocprofv3 --kernel-trace --stats` run and lists the kernels by
total time. `pmc` reads the *_counter_collection.csv of `rocprofv3 --pmc FETCH_SIZE` and `--pmc WRITE_SIZE`
runs (KiB per dispatch, summed over the counter's instances within a dispatch), averages each kernel's
dispatches, and scales FETCH_SIZE and WRITE_SIZE by what the same counters read for a plain device copy of
`--copy-bytes` bytes (the copy's bytes over the counter's bytes), giving HBM bytes per launch.
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import sys
from collections import defaultdict


def _one(pattern: str) -> str:
    found = sorted(glob.glob(pattern, recursive=True))
    if not found:
        raise SystemExit(f"no file matches {pattern}")
    return found[0]


def kernel_name(raw: str) -> str:
    """'void kspn::k_expand<0, 8>(unsigned int const*, ...)' -> 'kspn::k_expand<0, 8>'"""
    name = raw.strip()
    depth = 0
    for i, ch in enumerate(name):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            name = name[:i]
            break
    if name.startswith("void "):
        name = name[5:]
    return name.strip()


def read_stats(path: str):
    rows = []
    with open(path) as f:
        for r in csv.DictReader(f):
            rows.append({"name": kernel_name(r["Name"]), "calls": int(r["Calls"]),
                         "total_ns": float(r["TotalDurationNs"]), "avg_ns": float(r["AverageNs"]),
                         "pct": float(r["Percentage"])})
    merged = {}
    for r in rows:
        m = merged.setdefault(r["name"], {"name": r["name"], "calls": 0, "total_ns": 0.0, "pct": 0.0})
        m["calls"] += r["calls"]
        m["total_ns"] += r["total_ns"]
        m["pct"] += r["pct"]
    out = sorted(merged.values(), key=lambda m: -m["total_ns"])
    for m in out:
        m["avg_ns"] = m["total_ns"] / max(m["calls"], 1)
    return out


def read_pmc(path: str, counter: str):
    """per kernel: (dispatches, mean value per dispatch) of `counter`"""
    per_dispatch = defaultdict(float)
    names = {}
    with open(path) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] != counter:
                continue
            key = r.get("Dispatch_Id") or r.get("Correlation_Id")
            per_dispatch[key] += float(r["Counter_Value"])
            names[key] = kernel_name(r["Kernel_Name"])
    acc = defaultdict(lambda: [0, 0.0])
    for key, v in per_dispatch.items():
        a = acc[names[key]]
        a[0] += 1
        a[1] += v
    return {k: (n, s / n) for k, (n, s) in acc.items()}


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="mode", required=True)
    st = sub.add_parser("stats")
    st.add_argument("dir")
    st.add_argument("--title", required=True)
    st.add_argument("--command", action="append", required=True)
    st.add_argument("--top", type=int, default=25)
    pm = sub.add_parser("pmc")
    for k in ("fetch", "write", "copy_fetch", "copy_write"):
        pm.add_argument("--" + k.replace("_", "-"), required=True)
    pm.add_argument("--copy-bytes", type=float, required=True)
    pm.add_argument("--workload", required=True)
    pm.add_argument("--title", required=True)
    pm.add_argument("--command", action="append", required=True)
    pm.add_argument("--json", required=True)
    pm.add_argument("--top", type=int, default=25)
    a = ap.parse_args(argv)
    w = sys.stdout.write
    if a.mode == "stats":
        rows = read_stats(_one(os.path.join(a.dir, "**", "*kernel_stats.csv")))
        total = sum(r["total_ns"] for r in rows)
        w(f"# {a.title}\n\nCommand:\n\n```\n" + "\n".join(a.command) + "\n```\n\n")
        w(f"All kernels together: {total / 1e6:.3f} ms over {sum(r['calls'] for r in rows)} launches.\n\n")
        w("| kernel | calls | total ms | average us | share |\n|---|---:|---:|---:|---:|\n")
        for r in rows[: a.top]:
            w(f"| `{r['name']}` | {r['calls']} | {r['total_ns'] / 1e6:.3f} | {r['avg_ns'] / 1e3:.2f} | "
              f"{100.0 * r['total_ns'] / total:.1f} % |\n")
        return
    fetch = read_pmc(_one(os.path.join(a.fetch, "**", "*counter_collection.csv")), "FETCH_SIZE")
    write = read_pmc(_one(os.path.join(a.write, "**", "*counter_collection.csv")), "WRITE_SIZE")
    cf = read_pmc(_one(os.path.join(a.copy_fetch, "**", "*counter_collection.csv")), "FETCH_SIZE")
    cw = read_pmc(_one(os.path.join(a.copy_write, "**", "*counter_collection.csv")), "WRITE_SIZE")
    # the copy: the kernel that read the most in its calibration run
    copy_fetch_kib = max(v for _, v in cf.values())
    copy_write_kib = max(v for _, v in cw.values())
    fs = a.copy_bytes / (copy_fetch_kib * 1024.0)
    ws = a.copy_bytes / (copy_write_kib * 1024.0)
    kernels = {}
    for name in set(fetch) | set(write):
        nf, vf = fetch.get(name, (0, 0.0))
        nw, vw = write.get(name, (0, 0.0))
        kernels[name] = {"launches": max(nf, nw), "fetch_kib": vf, "write_kib": vw,
                         "bytes_per_launch": vf * 1024.0 * fs + vw * 1024.0 * ws}
    order = sorted(kernels, key=lambda k: -kernels[k]["bytes_per_launch"] * kernels[k]["launches"])
    with open(a.json, "w") as f:
        json.dump({"workload": a.workload, "fetch_scale": fs, "write_scale": ws, "copy_bytes": a.copy_bytes,
                   "per": "dispatch (mean over the kernel's dispatches)", "kernels": kernels}, f, indent=1)
        f.write("\n")
    w(f"# {a.title}\n\nCommands (each counter in a run of its own; counters are KiB per dispatch):\n\n```\n"
      + "\n".join(a.command) + "\n```\n\n")
    w(f"Calibration: a plain device copy of {a.copy_bytes:.0f} bytes reads FETCH_SIZE = {copy_fetch_kib:.1f} KiB "
      f"and WRITE_SIZE = {copy_write_kib:.1f} KiB per dispatch, so FETCH_SIZE is scaled by {fs:.3f} and "
      f"WRITE_SIZE by {ws:.3f} below. Values are means over each kernel's dispatches.\n\n")
    w("| kernel | dispatches | FETCH KiB per dispatch | WRITE KiB per dispatch | HBM MB per dispatch (scaled) | "
      "HBM MB, all dispatches |\n|---|---:|---:|---:|---:|---:|\n")
    for name in order[: a.top]:
        k = kernels[name]
        w(f"| `{name}` | {k['launches']} | {k['fetch_kib']:.1f} | {k['write_kib']:.1f} | "
          f"{k['bytes_per_launch'] / 1e6:.3f} | {k['bytes_per_launch'] * k['launches'] / 1e6:.1f} |\n")


if __name__ == "__main__":
    main()
