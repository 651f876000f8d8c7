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
  python -m gkmsvm_amd.rocpd_summary traffic FETCH_DIR WRITE_DIR COPY_FETCH_DIR
          COPY_WRITE_DIR [skip] [VALU_DIR]
      HBM bytes per dispatch of every k_* kernel: the FETCH_SIZE and
      WRITE_SIZE means (KiB per dispatch, one --pmc run each) scaled by what
      the same counters read for device copies of COPY_BYTES (256 MiB; the
      __amd_rocclr_copyBuffer dispatches of the COPY runs), and with VALU_DIR
      (an SQ_INSTS_VALU run) the VALU wave instructions per dispatch.  The
      "TRAFFIC <kernel> <bytes>" and "VALU <kernel> <instructions>" lines are
      what bench.py reads.
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


COPY_BYTES = 256 << 20


def base_name(kernel_name: str) -> str:
    """A kernel's name without return type, template arguments and
    parameters ("void k_index_tile<4, 1, false>(IndexArgs)" -> "k_index_tile")."""
    name = kernel_name.split("(")[0].split("<")[0].strip()
    return name.split()[-1] if name else name


def _means(dir_path: str, counter: str, skip=0, match: str = "") -> dict:
    """kernel base name -> (dispatches, mean value per dispatch) of one
    counter (template instances of one kernel pooled)."""
    n_skip = _skips(skip)
    per = defaultdict(list)
    for f in glob.glob(os.path.join(dir_path, "**", "*counter_collection.csv"), recursive=True):
        with open(f) as fh:
            for r in csv.DictReader(fh):
                name = r.get("Kernel_Name", "")
                if r["Counter_Name"] != counter or (match and match not in name):
                    continue
                per[base_name(name)].append((_dispatch_order(r), float(r["Counter_Value"])))
    out = {}
    for name, vs in per.items():
        vals = [v for _, v in sorted(vs)][n_skip(name):]
        if vals:
            out[name] = (len(vals), sum(vals) / len(vals))
    return out


def traffic(fetch_dir, write_dir, copy_fetch_dir, copy_write_dir, skip=0, valu_dir=None) -> str:
    cf = _means(copy_fetch_dir, "FETCH_SIZE", 0, "copyBuffer")
    cw = _means(copy_write_dir, "WRITE_SIZE", 0, "copyBuffer")
    if not cf or not cw:
        raise SystemExit("no __amd_rocclr_copyBuffer dispatches in the calibration runs")
    kf = COPY_BYTES / (1024.0 * max(v for _, v in cf.values()))
    kw = COPY_BYTES / (1024.0 * max(v for _, v in cw.values()))
    fetch = _means(fetch_dir, "FETCH_SIZE", skip, "k_")
    write = _means(write_dir, "WRITE_SIZE", skip, "k_")
    valu = _means(valu_dir, "SQ_INSTS_VALU", skip, "k_") if valu_dir else {}
    lines = [f"CALIBRATION FETCH_SIZE x{kf:.3f} WRITE_SIZE x{kw:.3f} (device copies of "
             f"{COPY_BYTES} bytes: FETCH_SIZE {max(v for _, v in cf.values()):.1f} KiB, "
             f"WRITE_SIZE {max(v for _, v in cw.values()):.1f} KiB per dispatch)",
             f"{'kernel':44s} {'dispatches':>10s} {'FETCH_KiB':>14s} {'WRITE_KiB':>14s} "
             f"{'HBM_bytes':>16s}"]
    rows = []
    for name in sorted(set(fetch) & set(write)):
        n, f_kib = fetch[name]
        _, w_kib = write[name]
        b = f_kib * 1024.0 * kf + w_kib * 1024.0 * kw
        rows.append((name, b))
        lines.append(f"{name:44s} {n:10d} {f_kib:14.1f} {w_kib:14.1f} {b:16.4g}")
    for name, b in rows:
        lines.append(f"TRAFFIC {name} {b:.6g}")
    for name, (n, v) in sorted(valu.items()):
        lines.append(f"VALU {name} {v:.6g}")
    return "\n".join(lines)


def read_profile(path: str) -> dict:
    """{"traffic": {kernel: bytes per dispatch}, "valu": {kernel: VALU wave
    instructions per dispatch}} from a profile summary's TRAFFIC / VALU
    lines (kernels by base name)."""
    out = {"traffic": {}, "valu": {}}
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) == 3 and parts[0] in ("TRAFFIC", "VALU"):
                out[parts[0].lower()][parts[1]] = float(parts[2])
    return out


if __name__ == "__main__":
    kind, path = sys.argv[1], sys.argv[2]
    if kind == "trace":
        print(trace(path, int(sys.argv[3]) if len(sys.argv) > 3 else 15))
    elif kind == "trace-csv":
        print(trace_csv(path, int(sys.argv[3]) if len(sys.argv) > 3 else 15,
                        sys.argv[4] if len(sys.argv) > 4 else 0))
    elif kind == "traffic":
        print(traffic(path, sys.argv[3], sys.argv[4], sys.argv[5],
                      sys.argv[6] if len(sys.argv) > 6 else 0,
                      sys.argv[7] if len(sys.argv) > 7 else None))
    else:
        print(pmc(path, sys.argv[3] if len(sys.argv) > 3 else "",
                  sys.argv[4] if len(sys.argv) > 4 else 0))
