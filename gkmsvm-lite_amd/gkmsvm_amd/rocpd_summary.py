"""Text summaries of rocprofv3 output, for the committed profiles.

  python -m gkmsvm_amd.rocpd_summary trace <results.db> [top]
      per-kernel calls, total and mean duration, share (kernel-trace database)
  python -m gkmsvm_amd.rocpd_summary pmc <dir> [kernel-substring]
      per-kernel mean of each counter per dispatch (counter_collection.csv)
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


def pmc(dir_path: str, match: str = "") -> str:
    files = glob.glob(os.path.join(dir_path, "**", "*counter_collection.csv"), recursive=True)
    acc = defaultdict(list)
    for f in files:
        with open(f) as fh:
            for r in csv.DictReader(fh):
                name = r.get("Kernel_Name", "")
                if match and match not in name:
                    continue
                acc[(name.split("(")[0][:58], r["Counter_Name"])].append(float(r["Counter_Value"]))
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
    else:
        print(pmc(path, sys.argv[3] if len(sys.argv) > 3 else ""))
