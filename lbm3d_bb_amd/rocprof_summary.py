"""Summaries of rocprofv3 output for the step kernel (text for profiles/).

  python -m lbm3d_bb_amd.rocprof_summary stats DIR
      per-kernel table from DIR/**/*kernel_stats.csv
  python -m lbm3d_bb_amd.rocprof_summary traffic --fetch DIR --write DIR \\
          --kernel k_step --copy-bytes B --out FILE.json
      mean FETCH_SIZE / WRITE_SIZE per dispatch (KiB) of the kernels whose
      name contains --kernel, scaled by what the same counters read for the
      plain device-to-device copy (__amd_rocclr_copyBuffer) of B bytes that
      bench.py --copy-probe runs in the same process.

The copy reads B and writes B bytes, so its counters give the bytes per
counted KiB; the scaled step-kernel figures are HBM bytes per launch.
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import sys
from collections import defaultdict

COPY_KERNEL = "__amd_rocclr_copyBuffer"


def _rows(directory: str, suffix: str) -> list[dict]:
    files = sorted(glob.glob(os.path.join(directory, "**", f"*{suffix}"), recursive=True))
    if not files:
        raise FileNotFoundError(f"no *{suffix} under {directory}")
    out = []
    for f in files:
        with open(f, newline="", encoding="utf-8") as fh:
            out.extend(csv.DictReader(fh))
    return out


def kernel_stats(directory: str) -> list[dict]:
    rows = _rows(directory, "kernel_stats.csv")
    return [{"name": r["Name"], "calls": int(r["Calls"]), "total_ns": float(r["TotalDurationNs"]),
             "avg_ns": float(r["AverageNs"]), "pct": float(r["Percentage"]), "min_ns": float(r["MinNs"]),
             "max_ns": float(r["MaxNs"])} for r in rows]


def counter_means(directory: str, counter: str) -> dict[str, tuple[float, int]]:
    """Mean counter value per dispatch, by kernel name: {name: (mean, dispatches)}.
    The copy kernel is keyed by its largest grid only: the probe's copies are
    the largest the process runs (smaller ones are the library's own
    host/device transfers)."""
    rows = _rows(directory, "counter_collection.csv")
    acc: dict[str, list[float]] = defaultdict(list)
    copy_grid = max((int(r["Grid_Size"]) for r in rows
                     if r.get("Counter_Name") == counter and COPY_KERNEL in r["Kernel_Name"]), default=-1)
    for r in rows:
        if r.get("Counter_Name") != counter:
            continue
        if COPY_KERNEL in r["Kernel_Name"] and int(r["Grid_Size"]) != copy_grid:
            continue
        acc[r["Kernel_Name"]].append(float(r["Counter_Value"]))
    return {k: (sum(v) / len(v), len(v)) for k, v in acc.items()}


def _match(means: dict, key: str) -> tuple[float, int]:
    hits = [(n, v) for n, v in means.items() if key in n]
    if not hits:
        raise KeyError(f"no kernel matching {key!r} among {list(means)[:8]}")
    tot = sum(v[0] * v[1] for _, v in hits)
    cnt = sum(v[1] for _, v in hits)
    return tot / cnt, cnt


def traffic(fetch_dir: str, write_dir: str, kernel: str, copy_bytes: float) -> dict:
    f = counter_means(fetch_dir, "FETCH_SIZE")
    w = counter_means(write_dir, "WRITE_SIZE")
    kf, nf = _match(f, kernel)
    kw, nw = _match(w, kernel)
    cf, _ = _match(f, COPY_KERNEL)
    cw, _ = _match(w, COPY_KERNEL)
    fetch_scale = copy_bytes / cf   # bytes per counted KiB of FETCH_SIZE
    write_scale = copy_bytes / cw
    fetch_b = kf * fetch_scale
    write_b = kw * write_scale
    return {"kernel": kernel, "dispatches": {"fetch": nf, "write": nw},
            "fetch_kib_per_dispatch": kf, "write_kib_per_dispatch": kw,
            "copy_bytes": copy_bytes, "copy_fetch_kib": cf, "copy_write_kib": cw,
            "fetch_bytes_per_kib": fetch_scale, "write_bytes_per_kib": write_scale,
            "fetch_bytes_per_launch": fetch_b, "write_bytes_per_launch": write_b,
            "bytes_per_launch": fetch_b + write_b}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("dir")
    s.add_argument("--top", type=int, default=12)
    t = sub.add_parser("traffic")
    t.add_argument("--fetch", required=True)
    t.add_argument("--write", required=True)
    t.add_argument("--kernel", default="k_step")
    t.add_argument("--copy-bytes", type=float, required=True)
    t.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    if a.cmd == "stats":
        rows = sorted(kernel_stats(a.dir), key=lambda r: -r["total_ns"])[: a.top]
        print(f"{'calls':>7} {'avg_us':>10} {'min_us':>10} {'max_us':>10} {'pct':>6}  kernel")
        for r in rows:
            name = r["name"] if len(r["name"]) < 110 else r["name"][:107] + "..."
            print(f"{r['calls']:7d} {r['avg_ns'] / 1e3:10.2f} {r['min_ns'] / 1e3:10.2f} {r['max_ns'] / 1e3:10.2f} "
                  f"{r['pct']:6.2f}  {name}")
        return 0
    res = traffic(a.fetch, a.write, a.kernel, a.copy_bytes)
    js = json.dumps(res, indent=1)
    print(js)
    if a.out:
        with open(a.out, "w", encoding="utf-8") as fh:
            fh.write(js + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
