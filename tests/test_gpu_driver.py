"""GPU tests of the reference-compatible drivers: the CLI/CSV shim and _time_one, checked
against the reference's recorded checksums the way its own validators do
(experiments/h100/tools/validate_cuda_vs_registry.py:45-106, validate_vs_registry.py:182-196)."""

from __future__ import annotations

import csv
import json
from pathlib import Path

import pytest

from zkduel_hip.cli import main as cli_main
from zkduel_hip.driver import time_one

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
REGISTRY = {(g["workload"], g["bit_width"], g["rounds"], g["challenge_mode"]): g["checksum_x2"]
            for g in json.loads((GOLDEN / "registry_checksums.json").read_text())}


def test_cli_csv_matches_registry_goldens(tmp_path):
    out = tmp_path / "hip.csv"
    for mode in ("fixed", "sha3"):
        rc = cli_main(["--workloads", "poly_abc", "zkphire_8", "zk_opencheck", "--bit-widths", "32", "256",
                       "--rounds", "10", "--warmups", "0", "--repeats", "1", "--challenge-mode", mode,
                       "--point-mode", "specialized", "--out", str(out)])
        assert rc == 0
        rows = list(csv.DictReader(out.open()))
        assert len(rows) == 6
        for row in rows:
            key = (row["workload"], int(row["bit_width"]), int(row["rounds"]), mode)
            assert int(row["checksum"]) == REGISTRY[key], key  # 1 + 0 + 1 runs = 2x one run
            assert row["status"] == "ok"


def test_time_one_row_and_checksum_multiplicity():
    row = time_one("zk_vanilla_permcheck_hp", 128, 10, warmups=2, repeats=3, seed=1,
                   timing_mode="whole-run", challenge_mode="sha3", layout="element", point_mode="specialized")
    one = REGISTRY[("zk_vanilla_permcheck_hp", 128, 10, "sha3")] // 2
    assert row["checksum"] == 6 * one  # first + 2 warmups + 3 repeats
    assert row["vars"] == 11 and row["degree"] == 5 and row["N"] == 1024
    assert row["round_polys"].shape == (10, 6, 4)
