"""Summaries of rocprofv3 output databases (the SQLite ``*_results.db`` files
that ``rocprofv3 -d DIR -o run`` writes): per-kernel time table of a
``--kernel-trace --stats`` run, per-dispatch means of ``--pmc`` counters,
and the HBM traffic of one kernel from a FETCH_SIZE run and a WRITE_SIZE run
calibrated on a plain copy of known size.

    python rocprof_summary.py trace DIR/run_results.db
    python rocprof_summary.py pmc DIR/run_results.db KERNEL_SUBSTRING
    python rocprof_summary.py traffic FETCH_DB WRITE_DB COPY_FETCH_DB COPY_WRITE_DB COPY_BYTES KERNEL
"""
from __future__ import annotations

import json
import sqlite3
import sys


def kernel_table(db: str, limit: int = 12) -> list[dict]:
    """Rows of the top_kernels view: name, calls, total ms, average us, percent."""
    con = sqlite3.connect(db)
    rows = con.execute("select name, total_calls, total_duration, average, percentage from top_kernels "
                       "order by total_duration desc limit ?", (limit,)).fetchall()
    # durations in the view are microseconds
    return [dict(name=r[0], calls=r[1], total_ms=r[2] / 1e3, avg_us=r[3], percent=r[4]) for r in rows]


def durations_us(db: str, kernel: str) -> list[float]:
    con = sqlite3.connect(db)
    return [r[0] / 1e3 for r in con.execute("select duration from kernels where name like ? order by start",
                                            (f"%{kernel}%",))]


def pmc_per_dispatch(db: str, kernel: str) -> dict[str, list[float]]:
    """Counter values summed over instances, one entry per dispatch of kernels whose name contains kernel."""
    con = sqlite3.connect(db)
    per: dict[tuple[int, str], float] = {}
    for name, counter, value, disp in con.execute(
            "select kernel_name, counter_name, value, dispatch_id from counters_collection"):
        if kernel in name:
            per[(disp, counter)] = per.get((disp, counter), 0.0) + value
    out: dict[str, list[float]] = {}
    for (disp, counter), v in sorted(per.items()):
        out.setdefault(counter, []).append(v)
    return out


def mean(xs):
    return sum(xs) / len(xs) if xs else float("nan")


def traffic(fetch_db, write_db, copy_fetch_db, copy_write_db, copy_bytes: float, kernel: str,
            copy_kernel: str = "copyBuffer") -> dict:
    """HBM bytes per dispatch of kernel: mean FETCH_SIZE and WRITE_SIZE (KiB per
    dispatch), each scaled by bytes / (KiB * 1024) the same counter reports for
    a copy kernel that reads and writes copy_bytes per dispatch."""
    f = mean(pmc_per_dispatch(fetch_db, kernel)["FETCH_SIZE"])
    w = mean(pmc_per_dispatch(write_db, kernel)["WRITE_SIZE"])
    cf = mean(pmc_per_dispatch(copy_fetch_db, copy_kernel)["FETCH_SIZE"])
    cw = mean(pmc_per_dispatch(copy_write_db, copy_kernel)["WRITE_SIZE"])
    kf, kw = copy_bytes / (cf * 1024.0), copy_bytes / (cw * 1024.0)
    return dict(kernel=kernel, fetch_kib=f, write_kib=w, copy_fetch_kib=cf, copy_write_kib=cw,
                fetch_scale=kf, write_scale=kw, bytes_per_launch=f * 1024.0 * kf + w * 1024.0 * kw,
                raw_bytes_per_launch=(f + w) * 1024.0)


def main(argv):
    cmd = argv[0]
    if cmd == "trace":
        for r in kernel_table(argv[1]):
            print(f"{r['calls']:6d} {r['total_ms']:10.1f} ms {r['avg_us']:10.1f} us {r['percent']:5.1f} %  {r['name'][:90]}")
    elif cmd == "pmc":
        for c, v in pmc_per_dispatch(argv[1], argv[2]).items():
            print(f"{c}: dispatches {len(v)} mean {mean(v):.6g}")
    elif cmd == "traffic":
        print(json.dumps(traffic(argv[1], argv[2], argv[3], argv[4], float(argv[5]), argv[6]), indent=1))
    else:
        raise SystemExit(__doc__)


if __name__ == "__main__":
    main(sys.argv[1:])
