"""Reduce rocprofv3 output to the per-kernel tables kept under profiles/.

    python spgemm-hpc-thesis_amd/py/pmc_summary.py counters <counter_collection.csv> [--kernel SUBSTR]
        [--dispatch last|all|median|N] [--segment MARKER:K] [--factor COUNTER=F ...]
    python spgemm-hpc-thesis_amd/py/pmc_summary.py trace <kernel_stats.csv> [--top N]
    python spgemm-hpc-thesis_amd/py/pmc_summary.py timeline <kernel_trace.csv> --segment MARKER:K [--min-us U]

counters: one row per (kernel, counter) with the value per dispatch of the
  chosen dispatches (`last`: the last dispatch of each kernel, which in a
  bench run is the last timed product; `all`: the mean over every dispatch;
  `median`: the lower median over every dispatch; N: the N-th dispatch,
  0-based; negative N counts from the end).  --segment MARKER:K first keeps
  only the dispatches of segment K (0-based, negative from the end) of the
  dispatch sequence cut before every dispatch of the kernel MARKER (for
  example one product's pipeline, which starts with k_zero_control), and
  reports each kernel's sum over its dispatches in that segment.  Counters
  with a calibration factor
  (--factor FETCH_SIZE=2.0, from a plain device copy of known size) are also
  printed in bytes (KiB x 1024 x factor).  Derived ratios are printed when
  their counters are present: SQ_LDS_BANK_CONFLICT / SQ_ACTIVE_INST_LDS and
  SQ_WAIT_ANY / SQ_WAVE_CYCLES.
trace: the --kernel-trace --stats summary (calls, average and total
  milliseconds, share) as a markdown table.
timeline: the dispatches of one segment of the --kernel-trace record (cut as
  for counters, by the dispatches of MARKER), in start order, with start and
  end in milliseconds from the segment's first start, so that kernels that
  overlap on two streams and the critical path are visible; dispatches
  shorter than --min-us microseconds are left out.
"""
from __future__ import annotations

import argparse
import csv
from collections import OrderedDict, defaultdict


def _short(name: str) -> str:
    """Kernel name without its namespace and argument list."""
    n = name
    for prefix in ("void ", "spg::(anonymous namespace)::", "(anonymous namespace)::", "spg::dev::", "spg::"):
        n = n.replace(prefix, "")
    return n.split("(")[0].strip()


def counters(path, kernel=None, dispatch="last", factors=None, segment=None):
    factors = factors or {}
    rows = defaultdict(lambda: defaultdict(list))  # kernel -> counter -> [(dispatch id, value)]
    with open(path, newline="") as f:
        recs = list(csv.DictReader(f))
    if segment:
        marker, k = segment.rsplit(":", 1)
        starts = sorted({int(r["Dispatch_Id"]) for r in recs if marker in _short(r["Kernel_Name"])})
        lo = starts[int(k)]
        nxt = [d for d in starts if d > lo]
        hi = nxt[0] if nxt else float("inf")
        recs = [r for r in recs if lo <= int(r["Dispatch_Id"]) < hi]
        dispatch = "sum"
    for r in recs:
        k = _short(r["Kernel_Name"])
        if kernel and kernel not in k:
            continue
        rows[k][r["Counter_Name"]].append((int(r["Dispatch_Id"]), float(r["Counter_Value"])))
    out = []
    for k, cs in rows.items():
        vals = OrderedDict()
        ndisp = 0
        for c, lst in cs.items():
            lst.sort()
            ndisp = max(ndisp, len(lst))
            if dispatch == "last":
                v = lst[-1][1]
            elif dispatch == "all":
                v = sum(x for _, x in lst) / len(lst)
            elif dispatch == "sum":
                v = sum(x for _, x in lst)
            elif dispatch == "median":
                xs = sorted(x for _, x in lst)
                v = xs[(len(xs) - 1) // 2]
            else:
                v = lst[int(dispatch)][1]
            vals[c] = v
        out.append((k, ndisp, vals))
    lines = []
    for k, ndisp, vals in out:
        what = f"segment {segment}, summed" if segment else dispatch
        lines.append(f"### `{k}` ({ndisp} dispatches; {what})")
        lines.append("")
        lines.append("| counter | per dispatch | bytes |")
        lines.append("|---|---|---|")
        for c, v in vals.items():
            b = f"{v * 1024 * factors[c]:,.0f}" if c in factors else ""
            lines.append(f"| {c} | {v:,.1f} | {b} |")
        if "SQ_LDS_BANK_CONFLICT" in vals and vals.get("SQ_ACTIVE_INST_LDS"):
            lines.append(f"| bank-conflict cycles / LDS-active cycles | "
                         f"{vals['SQ_LDS_BANK_CONFLICT'] / vals['SQ_ACTIVE_INST_LDS']:.3f} | |")
        if "SQ_WAIT_ANY" in vals and vals.get("SQ_WAVE_CYCLES"):
            lines.append(f"| wait cycles / wave cycles | {vals['SQ_WAIT_ANY'] / vals['SQ_WAVE_CYCLES']:.3f} | |")
        lines.append("")
    return "\n".join(lines)


def trace(path, top=20):
    with open(path, newline="") as f:
        rs = list(csv.DictReader(f))
    rs.sort(key=lambda r: -float(r["TotalDurationNs"]))
    lines = ["| kernel | calls | avg ms | total ms | % |", "|---|---|---|---|---|"]
    for r in rs[:top]:
        lines.append(f"| `{_short(r['Name'])}` | {r['Calls']} | {float(r['AverageNs']) / 1e6:.4f} | "
                     f"{float(r['TotalDurationNs']) / 1e6:.3f} | {float(r['Percentage']):.2f} |")
    return "\n".join(lines)


def timeline(path, segment, min_us=0.0):
    with open(path, newline="") as f:
        recs = list(csv.DictReader(f))
    recs.sort(key=lambda r: int(r["Start_Timestamp"]))
    marker, k = segment.rsplit(":", 1)
    starts = [i for i, r in enumerate(recs) if marker in _short(r["Kernel_Name"])]
    lo = starts[int(k)]
    nxt = [i for i in starts if i > lo]
    seg = recs[lo:nxt[0]] if nxt else recs[lo:]
    t0 = int(seg[0]["Start_Timestamp"])
    t_end = max(int(r["End_Timestamp"]) for r in seg)
    lines = [f"segment {segment}: {len(seg)} dispatches, first start to last end {(t_end - t0) / 1e6:.3f} ms", "",
             "| kernel | start ms | end ms | ms |", "|---|---|---|---|"]
    for r in seg:
        a, b = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
        if (b - a) / 1e3 < min_us:
            continue
        lines.append(f"| `{_short(r['Kernel_Name'])}` | {(a - t0) / 1e6:.3f} | {(b - t0) / 1e6:.3f} | {(b - a) / 1e6:.3f} |")
    return "\n".join(lines)


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="mode", required=True)
    c = sub.add_parser("counters")
    c.add_argument("csv")
    c.add_argument("--kernel", default=None)
    c.add_argument("--dispatch", default="last")
    c.add_argument("--segment", default=None)
    c.add_argument("--factor", action="append", default=[])
    t = sub.add_parser("trace")
    t.add_argument("csv")
    t.add_argument("--top", type=int, default=20)
    tl = sub.add_parser("timeline")
    tl.add_argument("csv")
    tl.add_argument("--segment", required=True)
    tl.add_argument("--min-us", type=float, default=0.0)
    a = ap.parse_args()
    if a.mode == "timeline":
        print(timeline(a.csv, a.segment, a.min_us))
    elif a.mode == "counters":
        fac = {}
        for item in a.factor:
            name, val = item.split("=")
            fac[name] = float(val)
        print(counters(a.csv, a.kernel, a.dispatch, fac, a.segment))
    else:
        print(trace(a.csv, a.top))


if __name__ == "__main__":
    main()
