"""Command-line twin of the reference's CUDA executable (cuda/src/field_sumcheck.cu:1091-1176,
1322-1366): same flags, same CSV header and rows, so drivers that spawn the executable and
join its CSV by (workload, rounds, bit_width) and checksum keep working
(benchmarks/field_sweep/compare_triton_cuda.py:108-136,
experiments/h100/tools/validate_cuda_vs_registry.py:66-91).

    python -m zkduel_hip.cli --workloads poly_abc --bit-widths 64 --rounds 10 16 \\
        --warmups 0 --repeats 1 --challenge-mode sha3 --out results.csv

Differences: the header and rows carry six more columns after the reference's 28
(proofs_per_s, p50_ms, p99_ms, algo_bytes, hbm_gbps, roofline_frac); with --timing-mode
round-sum, eval_median_ms is the device time of the pass kernels (each round's fold fused in)
and the tail, challenge_median_ms that of the finalize launches (sum of partials and device
transcript), and fold_median_ms is empty (no separate fold launch); --layout limb hands the
generated tables over limb-major ([table][limb][entry])
and the library transposes them once on the device, so the kernels always stream
element-major; --block is accepted and ignored (the pass grid is chosen per field); errors exit
with status 2 (bad flags) or 1 (runtime), as the reference.
"""

from __future__ import annotations

import csv
import sys
from pathlib import Path

from .driver import CSV_COLUMNS, EXTRA_COLUMNS, time_one
from .fields import FIELD_SPECS
from .workloads import ALIASES, SPEC_NAMES, ZK_SPEC_NAMES, canonical_workload_name


def _die(msg: str, code: int = 2) -> None:
    print(msg, file=sys.stderr)
    raise SystemExit(code)


def parse_args(argv: list[str]) -> dict:
    """Flag grammar of field_sumcheck.cu:1108-1176: list flags take values up to the next --."""
    a = {
        "rounds": [8, 10, 12, 14, 16], "bit_widths": [32, 64, 128, 256], "workloads": [],
        "warmups": 1, "repeats": 5, "seed": 1, "block": 128, "timing_mode": "whole-run",
        "challenge_mode": "fixed", "timer": "host", "layout": "element", "point_mode": "generic",
        "out": "results/cuda_field_sweep.csv", "zk_only": False,
    }
    i = 0

    def need(flag: str) -> str:
        nonlocal i
        if i + 1 >= len(argv):
            _die(f"{flag} requires a value")
        i += 1
        return argv[i]

    def many(cast) -> list:
        nonlocal i
        vals = []
        while i + 1 < len(argv) and not argv[i + 1].startswith("--"):
            i += 1
            vals.append(cast(argv[i]))
        return vals

    choices = {
        "--timing-mode": ("timing_mode", ("whole-run", "round-sum")),
        "--challenge-mode": ("challenge_mode", ("fixed", "sha3")),
        "--timer": ("timer", ("host", "events")),
        "--layout": ("layout", ("element", "limb")),
        "--point-mode": ("point_mode", ("generic", "specialized")),
    }
    while i < len(argv):
        f = argv[i]
        try:
            if f == "--rounds":
                a["rounds"] = many(int)
            elif f == "--bit-widths":
                a["bit_widths"] = many(int)
            elif f == "--workloads":
                a["workloads"] = many(str)
            elif f in ("--warmups", "--repeats", "--seed", "--block"):
                a[f[2:]] = int(need(f))
            elif f in choices:
                key, allowed = choices[f]
                v = need(f)
                if v not in allowed:
                    _die(f"{f} must be {' or '.join(allowed)}")
                a[key] = v
            elif f == "--zk-only":
                a["zk_only"] = True
            elif f == "--out":
                a["out"] = need(f)
            else:
                _die(f"unknown argument: {f}")
        except ValueError:
            _die(f"bad value for {f}")
        i += 1
    if a["block"] <= 0 or a["block"] & (a["block"] - 1):
        _die("--block must be a positive power of two")
    return a


def main(argv: list[str] | None = None) -> int:
    a = parse_args(sys.argv[1:] if argv is None else argv)
    if a["workloads"]:
        names = [canonical_workload_name(w) for w in a["workloads"]]
        known = set(SPEC_NAMES)
        order = [w for w in SPEC_NAMES if w in names]  # the executable's table order
        unknown = [w for w in names if w not in known]
        if unknown:
            _die(f"unknown workloads: {unknown}; known: {sorted(set(SPEC_NAMES) | set(ALIASES))}")
    else:
        order = list(ZK_SPEC_NAMES if a["zk_only"] else SPEC_NAMES)
    out = Path(a["out"])
    if out.parent:
        out.parent.mkdir(parents=True, exist_ok=True)
    with out.open("w", newline="") as fh:
        # the reference's 28 columns, then the throughput and roofline columns
        w = csv.DictWriter(fh, fieldnames=CSV_COLUMNS + EXTRA_COLUMNS, extrasaction="ignore")
        w.writeheader()
        for bw in sorted(FIELD_SPECS):
            if bw not in a["bit_widths"]:
                continue
            for name in order:
                for rounds in a["rounds"]:
                    row = time_one(name, bw, rounds, warmups=a["warmups"], repeats=a["repeats"],
                                   seed=a["seed"], timing_mode=a["timing_mode"],
                                   challenge_mode=a["challenge_mode"], layout=a["layout"],
                                   point_mode=a["point_mode"], timer=a["timer"])
                    w.writerow(row)
                    fh.flush()
                    print(f"{name} field={row['field']} rounds={rounds} first_ms={row['first_ms']} "
                          f"median_ms={row['median_ms']} checksum={row['checksum']}", flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
