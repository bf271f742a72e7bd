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


def test_limb_major_layout_matches_registry_goldens(tmp_path):
    """--layout limb (the reference's [table][limb][entry] tables, handed over limb-major and
    transposed on the device) reproduces the same recorded checksums."""
    out = tmp_path / "hip_limb.csv"
    rc = cli_main(["--workloads", "poly_abc_plus_de", "zk_complete_add_3", "--bit-widths", "64", "256",
                   "--rounds", "10", "--warmups", "0", "--repeats", "1", "--challenge-mode", "sha3",
                   "--layout", "limb", "--out", str(out)])
    assert rc == 0
    rows = list(csv.DictReader(out.open()))
    assert len(rows) == 4
    for row in rows:
        assert row["layout"] == "limb"
        key = (row["workload"], int(row["bit_width"]), int(row["rounds"]), "sha3")
        assert int(row["checksum"]) == REGISTRY[key], key


def test_limb_major_inputs_give_identical_proofs():
    """Host and device limb-major inputs give the bits of the element-major proof."""
    import ctypes

    import numpy as np

    from zkduel_hip import Composition, Context, Proof, get_spec

    ctx = Context(0)
    hip = ctypes.CDLL("libamdhip64.so")  # already loaded by the library; synchronous copies only
    comp = Composition.from_spec(get_spec("poly_abc_plus_de"))
    for bw in (32, 64, 128, 256):
        with Proof(ctx, bw, comp, 12, input="uniform", seed=21) as pr:
            pr.run("sha3")
            pr.wait()
            want = pr.result()
            limb_major = np.ascontiguousarray(pr.tables().transpose(0, 2, 1))
        with Proof(ctx, bw, comp, 12, input="host-mont-limb", data=limb_major) as pr:
            pr.run("sha3")
            pr.wait()
            got = pr.result()
        assert np.array_equal(got[0], want[0]) and got[2] == want[2]
        dev = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(dev), ctypes.c_size_t(limb_major.nbytes)) == 0
        try:
            assert hip.hipMemcpy(dev, limb_major.ctypes.data_as(ctypes.c_void_p),
                                 ctypes.c_size_t(limb_major.nbytes), 1) == 0  # host to device
            with Proof(ctx, bw, comp, 12, input="device-mont-limb", data=dev.value) as pr:
                pr.run("sha3")
                pr.wait()
                got = pr.result()
        finally:
            hip.hipFree(dev)
        assert np.array_equal(got[0], want[0]) and got[2] == want[2]
    ctx.close()


def test_round_sum_phase_columns_and_extra_columns(tmp_path):
    """--timing-mode round-sum fills eval (pass kernels with the fold fused in, and the tail)
    and challenge (finalize launches) from device events; fold has no launch of its own and
    is empty; the six extra columns follow the reference's 28."""
    from zkduel_hip.driver import CSV_COLUMNS, EXTRA_COLUMNS

    out = tmp_path / "rs.csv"
    rc = cli_main(["--workloads", "poly_abc_plus_de", "--bit-widths", "64", "256", "--rounds", "16",
                   "--warmups", "1", "--repeats", "3", "--timing-mode", "round-sum", "--challenge-mode", "sha3",
                   "--out", str(out)])
    assert rc == 0
    with out.open() as fh:
        header = fh.readline().strip().split(",")
    assert header == CSV_COLUMNS + EXTRA_COLUMNS
    rows = list(csv.DictReader(out.open()))
    assert len(rows) == 2
    for row in rows:
        ev, ch, total = float(row["eval_median_ms"]), float(row["challenge_median_ms"]), float(row["total_median_ms"])
        assert ev > 0 and ch > 0 and ev + ch <= total * 1.05, row
        assert row["fold_median_ms"] == ""
        assert float(row["proofs_per_s"]) > 0 and int(row["algo_bytes"]) > 0
        assert 0 < float(row["roofline_frac"]) < 1
    whole = time_one("poly_abc_plus_de", 64, 12, warmups=0, repeats=2, timing_mode="whole-run")
    assert whole["eval_median_ms"] == "0.000" and whole["fold_median_ms"] == "0.000"
