"""Summaries of rocprofv3 output for the profiles/ folder.

    python -m gfaz_amd.profile_summary stats <kernel_stats.csv> --cmd "<command>"
    python -m gfaz_amd.profile_summary traffic --fetch <dir> --write <dir> --cmd "<cmd>" \
        --copy-bytes N

`stats` turns a `rocprofv3 --kernel-trace --stats` kernel_stats.csv into a short
markdown table. `traffic` combines two `rocprofv3 --pmc` runs (FETCH_SIZE, then
WRITE_SIZE; KiB per dispatch) of `bench.py --pmc-pass`, whose first kernel is a
plain copy of --copy-bytes bytes: the counter values that copy reads give the
scale factors from counter units to bytes. The traffic summary embeds one JSON
line (`<!-- traffic-json: {...} -->`) that bench.py reads for `roofline.traffic`.
"""
from __future__ import annotations

import argparse
import collections
import csv
import glob
import json
import os
import re
import sys

TRAFFIC_TAG = "<!-- traffic-json:"


def short_name(kernel: str) -> str:
    m = re.search(r"k_(\w+?)(\(|<)", kernel)
    if m:
        return m.group(1)
    return kernel.split("(")[0].split("::")[-1][:40]


def stats_markdown(csv_path: str, cmd: str, top: int = 0) -> str:
    """Every kernel (top > 0: only the top kernels by total time)."""
    rows = list(csv.DictReader(open(csv_path)))
    agg = collections.OrderedDict()
    for r in rows:
        name = short_name(r["Name"])
        m = re.search(r"k_\w+<([^()]*)>\(", r["Name"])  # the kernel's own template arguments
        if m:
            name = f"{name}<{m.group(1).replace('u', '').replace(' ', '')}>"
        a = agg.setdefault(name, [0, 0.0, float("inf"), 0.0])
        a[0] += int(r["Calls"])
        a[1] += float(r["TotalDurationNs"])
        a[2] = min(a[2], float(r["MinNs"]))
        a[3] = max(a[3], float(r["MaxNs"]))
    total = sum(v[1] for v in agg.values()) or 1.0
    lines = [f"Command: `{cmd}`", "",
             "| kernel | calls | total ms | avg us | min us | max us | % |",
             "|---|---:|---:|---:|---:|---:|---:|"]
    ranked = sorted(agg.items(), key=lambda kv: -kv[1][1])
    for name, (calls, tot, mn, mx) in ranked[:top] if top > 0 else ranked:
        lines.append(f"| {name} | {calls} | {tot / 1e6:.3f} | {tot / calls / 1e3:.1f} | "
                     f"{mn / 1e3:.1f} | {mx / 1e3:.1f} | {100 * tot / total:.1f} |")
    return "\n".join(lines) + "\n"


def _counter_rows(d: str):
    files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise SystemExit(f"no counter_collection.csv under {d}")
    for f in files:
        yield from csv.DictReader(open(f))


def per_dispatch(d: str, counter: str):
    """{kernel short name: [value per dispatch, in dispatch order]}"""
    vals = collections.defaultdict(dict)
    for r in _counter_rows(d):
        if r["Counter_Name"] != counter:
            continue
        k = short_name(r["Kernel_Name"])
        did = int(r["Dispatch_Id"])
        vals[k][did] = vals[k].get(did, 0.0) + float(r["Counter_Value"])
    return {k: [v[i] for i in sorted(v)] for k, v in vals.items()}


def traffic_markdown(fetch_dir: str, write_dir: str, cmd_fetch: str, cmd_write: str,
                     copy_bytes: int, copy_kernel: str = "copy16") -> str:
    fetch = per_dispatch(fetch_dir, "FETCH_SIZE")
    write = per_dispatch(write_dir, "WRITE_SIZE")
    if copy_kernel not in fetch or copy_kernel not in write:
        raise SystemExit(f"calibration kernel {copy_kernel} missing from the counter runs")
    f_copy, w_copy = fetch[copy_kernel][0], write[copy_kernel][0]
    f_scale = copy_bytes / f_copy   # bytes per counter unit, reads
    w_scale = copy_bytes / w_copy   # bytes per counter unit, writes
    table = {}
    for k in sorted(set(fetch) | set(write)):
        if k == copy_kernel:
            continue
        fv, wv = fetch.get(k, []), write.get(k, [])
        n = max(len(fv), len(wv))
        fb = sum(fv) * f_scale / max(len(fv), 1)
        wb = sum(wv) * w_scale / max(len(wv), 1)
        per = [(fv[i] if i < len(fv) else 0.0) * f_scale + (wv[i] if i < len(wv) else 0.0) * w_scale
               for i in range(n)]
        table[k] = {"dispatches": n, "fetch_bytes_per_dispatch": fb,
                    "write_bytes_per_dispatch": wb, "bytes_per_dispatch": fb + wb,
                    "bytes": [round(b) for b in per]}  # each dispatch, in launch order
    meta = {"copy_bytes": copy_bytes, "copy_fetch_counter": f_copy, "copy_write_counter": w_copy,
            "fetch_scale": f_scale, "write_scale": w_scale, "kernels": table}
    lines = [f"FETCH_SIZE run: `{cmd_fetch}`", "", f"WRITE_SIZE run: `{cmd_write}`", "",
             f"Calibration: a plain copy of {copy_bytes} bytes reads FETCH_SIZE = {f_copy:.0f} "
             f"and WRITE_SIZE = {w_copy:.0f} (KiB units): {f_scale:.1f} bytes per FETCH unit, "
             f"{w_scale:.1f} bytes per WRITE unit.", "",
             "| kernel | dispatches | fetch MB / dispatch | write MB / dispatch | total MB / dispatch |",
             "|---|---:|---:|---:|---:|"]
    for k, v in sorted(table.items(), key=lambda kv: -kv[1]["bytes_per_dispatch"] * kv[1]["dispatches"]):
        lines.append(f"| {k} | {v['dispatches']} | {v['fetch_bytes_per_dispatch'] / 1e6:.2f} | "
                     f"{v['write_bytes_per_dispatch'] / 1e6:.2f} | {v['bytes_per_dispatch'] / 1e6:.2f} |")
    lines += ["", f"{TRAFFIC_TAG} {json.dumps(meta)} -->"]
    return "\n".join(lines) + "\n"


def load_traffic(path: str) -> dict | None:
    """The embedded JSON of a traffic summary, or None."""
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(TRAFFIC_TAG):
            return json.loads(line[len(TRAFFIC_TAG):].rsplit("-->", 1)[0])
    return None


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="mode", required=True)
    s = sub.add_parser("stats")
    s.add_argument("csv")
    s.add_argument("--cmd", required=True)
    s.add_argument("--top", type=int, default=0, help="0: every kernel")
    t = sub.add_parser("traffic")
    t.add_argument("--fetch", required=True)
    t.add_argument("--write", required=True)
    t.add_argument("--cmd-fetch", required=True)
    t.add_argument("--cmd-write", required=True)
    t.add_argument("--copy-bytes", type=int, default=1 << 30)
    a = ap.parse_args(argv)
    if a.mode == "stats":
        sys.stdout.write(stats_markdown(a.csv, a.cmd, a.top))
    else:
        sys.stdout.write(traffic_markdown(a.fetch, a.write, a.cmd_fetch, a.cmd_write,
                                          a.copy_bytes))
    return 0


if __name__ == "__main__":
    sys.exit(main())
