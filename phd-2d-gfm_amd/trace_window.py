"""Per-kernel summary of a rocprofv3 CSV output directory over a window of dispatches.

bench.py replays its timed steps eagerly at the end of a run (one flux-kernel launch per step), so the
last N dispatches of a kernel are exactly that window.  This prints, as markdown:
  - kernel trace (*kernel_trace.csv): for every kernel, dispatches, mean / min / max duration over its
    last N dispatches (or all of them with --all), and its register, LDS and scratch allocation;
  - counters (*counter_collection.csv): for every kernel and counter, the mean value per dispatch over
    its last N dispatches.

usage: python trace_window.py DIR [--last N] [--kernel REGEX] [--all]
"""
import argparse
import csv
import glob
import os
import re
import sys
from collections import OrderedDict, defaultdict


def _rows(d, suffix):
    out = []
    for p in sorted(glob.glob(os.path.join(d, "**", "*" + suffix), recursive=True)):
        with open(p, newline="") as f:
            out.extend(csv.DictReader(f))
    return out


def _short(name):
    """kernel name without the argument list"""
    depth, out = 0, []
    for ch in name:
        if ch == "(" and depth == 0:
            break
        depth += ch == "<"
        depth -= ch == ">"
        out.append(ch)
    return "".join(out).strip()


def kernel_table(d, last, pattern, use_all):
    rows = _rows(d, "kernel_trace.csv")
    by = OrderedDict()
    for r in sorted(rows, key=lambda r: int(r["Start_Timestamp"])):
        by.setdefault(_short(r["Kernel_Name"]), []).append(r)
    lines = ["| kernel | dispatches | window | mean us | min us | max us | VGPR | AGPR | SGPR | LDS B | scratch B |",
             "|---|---:|---:|---:|---:|---:|---:|---:|---:|---:|---:|"]
    for k, rs in by.items():
        if pattern and not re.search(pattern, k):
            continue
        w = rs if use_all or last <= 0 else rs[-last:]
        dur = [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3 for r in w]
        r0 = w[-1]
        lines.append("| `%s` | %d | %d | %.1f | %.1f | %.1f | %s | %s | %s | %s | %s |" % (
            k, len(rs), len(w), sum(dur) / len(dur), min(dur), max(dur), r0.get("VGPR_Count", ""),
            r0.get("Accum_VGPR_Count", ""), r0.get("SGPR_Count", ""), r0.get("LDS_Block_Size", ""),
            r0.get("Scratch_Size", "")))
    return lines


def counter_table(d, last, pattern, use_all):
    rows = _rows(d, "counter_collection.csv")
    if not rows:
        return []
    # dispatch -> kernel, counter -> value (a counter may come in several rows per dispatch: summed)
    disp = OrderedDict()
    for r in sorted(rows, key=lambda r: int(r["Dispatch_Id"])):
        e = disp.setdefault(int(r["Dispatch_Id"]), {"k": _short(r["Kernel_Name"]), "c": defaultdict(float)})
        e["c"][r["Counter_Name"]] += float(r["Counter_Value"])
    by = OrderedDict()
    for e in disp.values():
        by.setdefault(e["k"], []).append(e["c"])
    names = sorted({c for e in disp.values() for c in e["c"]})
    lines = ["| kernel | window | " + " | ".join(names) + " |", "|---|---:|" + "---:|" * len(names)]
    for k, cs in by.items():
        if pattern and not re.search(pattern, k):
            continue
        w = cs if use_all or last <= 0 else cs[-last:]
        vals = ["%.4g" % (sum(c.get(n, 0.0) for c in w) / len(w)) for n in names]
        lines.append("| `%s` | %d | %s |" % (k, len(w), " | ".join(vals)))
    return lines


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("dir")
    ap.add_argument("--last", type=int, default=0, help="last N dispatches of each kernel (0: all)")
    ap.add_argument("--kernel", default=None, help="regex on the kernel name")
    ap.add_argument("--all", action="store_true")
    a = ap.parse_args(argv)
    out = kernel_table(a.dir, a.last, a.kernel, a.all)
    if len(out) > 2:
        print("\n".join(out))
    ct = counter_table(a.dir, a.last, a.kernel, a.all)
    if ct:
        print()
        print("\n".join(ct))
    return 0


if __name__ == "__main__":
    sys.exit(main())
