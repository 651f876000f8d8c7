"""Text summaries of rocprofv3 output, for the committed profiles.

  python -m gkmsvm_amd.rocpd_summary trace <results.db> [top]
      per-kernel calls, total and mean duration, share (kernel-trace database)
  python -m gkmsvm_amd.rocpd_summary trace-csv <dir> [top] [skip]
      the same from kernel_trace.csv (--output-format csv), leaving out each
      kernel's first `skip` dispatches (a benchmark's warm-up launches):
      skip = N for every kernel, or "name=N,name=N" by kernel-name substring
  python -m gkmsvm_amd.rocpd_summary pmc <dir> [kernel-substring] [skip]
      per-kernel mean of each counter per dispatch (counter_collection.csv),
      leaving out each kernel's first `skip` dispatches
"""

from __future__ import annotations

import csv
import glob
import os
import sqlite3
import sys
from collections import defaultdict


def trace(db_path: str, top: int = 15) -> str:
    db = sqlite3.connect(db_path)
    rows = db.execute("select name, total_calls, total_duration, average, percentage "
                      "from top_kernels order by total_duration desc").fetchall()
    # the top_kernels view reports durations in microseconds
    lines = [f"{'kernel':58s} {'calls':>6s} {'total_ms':>10s} {'mean_us':>10s} {'share%':>7s}"]
    for name, calls, tot, avg, pct in rows[:top]:
        lines.append(f"{name.split('(')[0][:58]:58s} {calls:6d} {tot / 1e3:10.3f} "
                     f"{avg:10.1f} {pct:7.2f}")
    return "\n".join(lines)


def _skips(skip) -> callable:
    """skip: N (every kernel) or "sub=N,sub=N" (kernels whose name holds sub)."""
    if isinstance(skip, int) or str(skip).isdigit():
        n = int(skip)
        return lambda name: n
    table = [(k, int(v)) for k, v in (x.split("=") for x in str(skip).split(",") if x)]
    return lambda name: next((v for k, v in table if k in name), 0)


def _dispatch_order(r: dict) -> int:
    for key in ("Dispatch_Id", "Correlation_Id", "Start_Timestamp"):
        if r.get(key, "") not in ("", None):
            return int(float(r[key]))
    return 0


def trace_csv(dir_path: str, top: int = 15, skip=0) -> str:
    n_skip = _skips(skip)
    files = glob.glob(os.path.join(dir_path, "**", "*kernel_trace.csv"), recursive=True)
    per = defaultdict(list)
    for f in files:
        with open(f) as fh:
            for r in csv.DictReader(fh):
                per[r.get("Kernel_Name", "").split("(")[0][:58]].append(
                    (int(r["Start_Timestamp"]), int(r["End_Timestamp"]) - int(r["Start_Timestamp"])))
    rows = []
    for name, ds in per.items():
        ds = [d for _, d in sorted(ds)][n_skip(name):]
        if ds:
            rows.append((name, len(ds), sum(ds)))
    grand = sum(t for _, _, t in rows) or 1
    rows.sort(key=lambda x: -x[2])
    lines = [f"{'kernel':58s} {'calls':>6s} {'total_ms':>10s} {'mean_us':>10s} {'share%':>7s}"]
    for name, n, tot in rows[:top]:  # timestamps in nanoseconds
        lines.append(f"{name:58s} {n:6d} {tot / 1e6:10.3f} {tot / n / 1e3:10.1f} "
                     f"{100.0 * tot / grand:7.2f}")
    return "\n".join(lines)


def pmc(dir_path: str, match: str = "", skip=0) -> str:
    n_skip = _skips(skip)
    files = glob.glob(os.path.join(dir_path, "**", "*counter_collection.csv"), recursive=True)
    per = defaultdict(list)
    for f in files:
        with open(f) as fh:
            for r in csv.DictReader(fh):
                name = r.get("Kernel_Name", "")
                if match and match not in name:
                    continue
                per[(name.split("(")[0][:58], r["Counter_Name"])].append(
                    (_dispatch_order(r), float(r["Counter_Value"])))
    acc = {k: [v for _, v in sorted(vs)][n_skip(k[0]):] for k, vs in per.items()}
    acc = {k: v for k, v in acc.items() if v}
    lines = [f"{'kernel':58s} {'counter':>12s} {'dispatches':>10s} {'mean/dispatch':>16s} "
             f"{'median':>14s}"]
    for (name, ctr), vals in sorted(acc.items()):
        med = sorted(vals)[len(vals) // 2]
        lines.append(f"{name:58s} {ctr:>12s} {len(vals):10d} {sum(vals) / len(vals):16.1f} "
                     f"{med:14.1f}")
    return "\n".join(lines)


if __name__ == "__main__":
    kind, path = sys.argv[1], sys.argv[2]
    if kind == "trace":
        print(trace(path, int(sys.argv[3]) if len(sys.argv) > 3 else 15))
    elif kind == "trace-csv":
        print(trace_csv(path, int(sys.argv[3]) if len(sys.argv) > 3 else 15,
                        sys.argv[4] if len(sys.argv) > 4 else 0))
    else:
        print(pmc(path, sys.argv[3] if len(sys.argv) > 3 else "",
                  sys.argv[4] if len(sys.argv) > 4 else 0))
