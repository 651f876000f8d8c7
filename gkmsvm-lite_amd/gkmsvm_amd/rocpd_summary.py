"""Print a per-kernel summary (calls, total, mean, share) from a rocprofv3
results database (rocpd SQLite, e.g. gpurun_out/prof/run_results.db)."""
import sqlite3
import sys

db = sqlite3.connect(sys.argv[1])
rows = db.execute("select name, total_calls, total_duration, average, percentage "
                  "from top_kernels order by total_duration desc").fetchall()
print(f"{'kernel':60s} {'calls':>6s} {'total_ms':>10s} {'mean_us':>10s} {'share%':>7s}")
for name, calls, tot, avg, pct in rows[: int(sys.argv[2]) if len(sys.argv) > 2 else 15]:
    short = name.split("(")[0][:60]
    print(f"{short:60s} {calls:6d} {tot / 1e3:10.3f} {avg:10.1f} {pct:7.2f}")
