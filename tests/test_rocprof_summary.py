"""The rocprofv3 database summarizer on a small synthetic database (no GPU): launch statistics,
counters summed within a dispatch and averaged over dispatches, and recorded resources."""

import sqlite3
import sys
from pathlib import Path

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "sumcheck-triton_amd"))

from zkduel_hip import rocprof_summary as rs  # noqa: E402


def _db() -> sqlite3.Connection:
    db = sqlite3.connect(":memory:")
    db.execute("create table kernels (name text, start int, end int, dispatch_id int, arch_vgpr_count int, "
               "scratch_size int)")
    db.executemany("insert into kernels values (?, ?, ?, ?, ?, ?)", [
        ("pass", 0, 1000, 1, 98, 0), ("pass", 2000, 5000, 2, 98, 0), ("fin", 6000, 6500, 3, 40, 0)])
    db.execute("create table pmc_events (name text, dispatch_id int, counter_name text, counter_value real)")
    # two rows per dispatch (two hardware instances), to be summed within the dispatch
    db.executemany("insert into pmc_events values (?, ?, ?, ?)", [
        ("pass", 1, "FETCH_SIZE", 10.0), ("pass", 1, "FETCH_SIZE", 30.0),
        ("pass", 2, "FETCH_SIZE", 20.0), ("pass", 2, "FETCH_SIZE", 20.0),
        ("fin", 3, "FETCH_SIZE", 1.0), ("fin", 3, "FETCH_SIZE", 2.0),
        ("pass", 1, "WRITE_SIZE", 5.0)])
    return db


def test_kernel_stats_are_per_launch():
    rows = {r[0]: r[1:] for r in rs.kernel_stats(_db())}
    n, avg, mn, mx, tot = rows["pass"]
    assert (n, avg, mn, mx, tot) == (2, 2000.0, 1000, 3000, 4000)
    assert rows["fin"][0] == 1


def test_counters_sum_within_a_dispatch_then_average():
    per = rs.pmc_per_kernel(_db(), "FETCH_SIZE")
    assert per["pass"] == (2, 40.0)
    assert per["fin"] == (1, 3.0)
    assert rs.pmc_per_dispatch(_db(), "FETCH_SIZE", "pass") == [(1, 40.0), (2, 40.0)]


def test_resources_list_the_recorded_columns():
    have, rows = rs.kernel_resources(_db())
    assert have == ["arch_vgpr_count", "scratch_size"]
    assert rows == [("fin", 40, 0), ("pass", 98, 0)]
