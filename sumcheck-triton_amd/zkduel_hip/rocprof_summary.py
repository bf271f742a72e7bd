"""Summarize a rocprofv3 results database (``--kernel-trace`` and/or ``--pmc``) as text.

    python -m zkduel_hip.rocprof_summary <results.db> [--pmc COUNTER] [--match SUBSTRING] [--dispatches]

Prints per-kernel launch counts and durations (avg/min/max/total), and, with --pmc, the
counter's average value per dispatch of each kernel. Counters such as FETCH_SIZE and
WRITE_SIZE are in KiB per dispatch; dividing a kernel's value by the same counter's value
for a copy of known size (the library's calibration copy, ``calib_copy``, moves 2 GiB read
and 2 GiB written) turns them into bytes with the probe's own scale.
"""

from __future__ import annotations

import argparse
import sqlite3
from collections import defaultdict


def kernel_stats(db: sqlite3.Connection, match: str | None = None) -> list[tuple]:
    rows = db.execute(
        "select name, count(*), avg(end-start), min(end-start), max(end-start), sum(end-start) "
        "from kernels group by name order by sum(end-start) desc"
    ).fetchall()
    return [r for r in rows if match is None or match in r[0]]


def pmc_per_kernel(db: sqlite3.Connection, counter: str) -> dict[str, tuple[int, float]]:
    """{kernel: (dispatches, counter per dispatch averaged over them)}; a dispatch's rows (one
    per hardware instance, when the database keeps them apart) are summed first."""
    cols = [r[1] for r in db.execute("pragma table_info(pmc_events)")]
    name_col = "name" if "name" in cols else None
    counter_col = "counter_name" if "counter_name" in cols else ("pmc_name" if "pmc_name" in cols else None)
    value_col = "counter_value" if "counter_value" in cols else ("value" if "value" in cols else None)
    if not (name_col and counter_col and value_col and "dispatch_id" in cols):
        raise SystemExit(f"unexpected pmc_events columns: {cols}")
    acc: dict[str, list[float]] = defaultdict(list)
    for name, val in db.execute(
        f"select {name_col}, sum({value_col}) from pmc_events where {counter_col} = ? "
        f"group by dispatch_id", (counter,)
    ):
        acc[name].append(float(val))
    return {k: (len(v), sum(v) / len(v)) for k, v in acc.items()}


RESOURCE_COLUMNS = ("vgpr_count", "arch_vgpr_count", "accum_vgpr_count", "sgpr_count", "lds_size", "static_lds_size",
                    "group_segment_size", "scratch_size", "static_scratch_size", "private_segment_size")


def kernel_resources(db: sqlite3.Connection, match: str | None = None) -> tuple[list[str], list[tuple]]:
    """Per-kernel register, LDS and scratch sizes the trace recorded (whichever of
    RESOURCE_COLUMNS the database's kernels view has), one row per distinct kernel."""
    cols = [r[1] for r in db.execute("pragma table_info(kernels)")]
    have = [c for c in RESOURCE_COLUMNS if c in cols]
    if not have:
        return [], []
    rows = db.execute(f"select distinct name, {', '.join(have)} from kernels order by name").fetchall()
    return have, [r for r in rows if match is None or match in r[0]]


def pmc_per_dispatch(db: sqlite3.Connection, counter: str, match: str) -> list[tuple[int, float]]:
    """(dispatch id, counter value summed over the dispatch's records) of every dispatch of the
    kernels whose name contains ``match``, in dispatch order."""
    return [(int(d), float(v)) for d, v in db.execute(
        "select dispatch_id, sum(counter_value) from pmc_events where counter_name = ? and name like ? "
        "group by dispatch_id order by dispatch_id", (counter, f"%{match}%"))]


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("db")
    ap.add_argument("--pmc", action="append", default=[])
    ap.add_argument("--match", default=None)
    ap.add_argument("--dispatches", action="store_true",
                    help="with --pmc and --match: print the counter of every matching dispatch")
    ap.add_argument("--resources", action="store_true",
                    help="also print each kernel's registers, LDS and scratch as the trace recorded them")
    args = ap.parse_args()
    db = sqlite3.connect(args.db)
    print(f"{'kernel':70s} {'n':>5s} {'avg_us':>10s} {'min_us':>10s} {'max_us':>10s} {'total_ms':>10s}")
    for name, n, avg, mn, mx, tot in kernel_stats(db, args.match):
        print(f"{name[:70]:70s} {n:5d} {avg / 1e3:10.2f} {mn / 1e3:10.2f} {mx / 1e3:10.2f} {tot / 1e6:10.3f}")
    if args.resources:
        have, rows = kernel_resources(db, args.match)
        print("\n" + (f"{'kernel':70s} " + " ".join(f"{c:>20s}" for c in have) if have
                      else "(the trace recorded no register or scratch sizes)"))
        for r in rows:
            print(f"{r[0][:70]:70s} " + " ".join(f"{v!s:>20s}" for v in r[1:]))
    if args.dispatches and args.match:
        for counter in args.pmc:
            print(f"\n{counter} of each dispatch of kernels matching {args.match!r}")
            for d, v in pmc_per_dispatch(db, counter, args.match):
                print(f"{d:8d} {v:16.1f}")
        return 0
    for counter in args.pmc:
        print(f"\n{counter} per dispatch (average over dispatches)")
        for name, (n, avg) in sorted(pmc_per_kernel(db, counter).items(), key=lambda kv: -kv[1][1]):
            if args.match is None or args.match in name or "calib" in name:
                print(f"{name[:70]:70s} {n:5d} {avg:16.1f}")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
