"""Field-sweep driver: the reference's ``_time_one`` contract on the HIP library.

``time_one(workload, bit_width, rounds, warmups, repeats, seed, timing_mode, challenge_mode,
layout, point_mode)`` prepares one case, runs it ``1 + warmups + repeats`` times and returns
the reference's CSV row (benchmarks/field_sweep/triton_sweep.py:626-816; columns of
cuda/src/field_sumcheck.cu:1336), with the checksum summed over all runs as the reference
does (field_sumcheck.cu:1222-1285), plus ``round_polys`` of the last run and throughput
columns.

Reference-compatible inputs are the affine tables with seed ``seed + rounds``
(field_sumcheck.cu:1217-1219, triton_sweep.py:642).
"""

from __future__ import annotations

import statistics
import time

import numpy as np

from .fields import get_field
from .prover import Composition, Context, Proof, default_context
from .workloads import get_spec

CSV_COLUMNS = [
    "workload", "polynomial", "backend", "field", "arithmetic", "layout", "point_mode",
    "bit_width", "limbs", "modulus_hex", "rounds", "N", "vars", "degree", "terms", "warmups",
    "repeats", "timing_mode", "challenge_mode", "first_ms", "median_ms", "eval_median_ms",
    "challenge_median_ms", "fold_median_ms", "total_median_ms", "input_mib", "checksum", "status",
]
# Written after the reference's columns (joins that read rows by name ignore them): whole-proof
# throughput and latency percentiles over the repeats, the algorithmic HBM bytes of one proof,
# the rate they are moved at, and that rate over the 8 TB/s peak.
EXTRA_COLUMNS = ["proofs_per_s", "p50_ms", "p99_ms", "algo_bytes", "hbm_gbps", "roofline_frac"]
HBM_PEAK_GBPS = 8000.0


def algorithmic_bytes(num_tables: int, limbs: int, num_vars: int) -> int:
    """HBM bytes of one fused proof: each round streams its input once and, from round 1 on,
    writes the folded tables: T*W*(3*2^nv - 4) read + T*W*(2^nv - 2) written, W = 4L."""
    w = 4 * limbs
    n = 1 << num_vars
    return num_tables * w * (3 * n - 4) + num_tables * w * (n - 2)


def run_timed(proof: Proof, challenge_mode: str, timer: str = "host", challenges=None) -> float:
    """One whole proof, synchronized on both sides; ms by host clock or device events."""
    proof.ctx.sync()
    t0 = time.perf_counter()
    proof.run(challenge_mode, challenges)
    dev_ms = proof.wait()
    host_ms = (time.perf_counter() - t0) * 1000.0
    return dev_ms if timer == "events" else host_ms


def phase_ms(proof: Proof) -> tuple[float, float]:
    """(eval, challenge) device ms of the last profiled run, the reference's round-sum phases
    (field_sumcheck.cu:1031-1084): eval = every round's pass launches, whose fold of the
    previous round is fused in (point-parallel rounds: their fold and evaluation launches),
    plus the tail kernel, which runs the last rounds' eval, fold and challenge in one launch;
    challenge = the finalize launches (partials summed, g_j stored, r_j derived)."""
    passes, fins, tail = proof.round_times()
    return sum(passes) + tail, sum(fins)


def time_one(
    workload: str,
    bit_width: int,
    rounds: int,
    warmups: int = 1,
    repeats: int = 5,
    seed: int = 1,
    timing_mode: str = "whole-run",
    challenge_mode: str = "fixed",
    layout: str = "element",
    point_mode: str = "specialized",
    *,
    input_mode: str = "affine",
    timer: str = "host",
    ctx: Context | None = None,
) -> dict:
    if timing_mode not in ("whole-run", "round-sum"):
        raise ValueError("timing_mode must be whole-run or round-sum")
    if challenge_mode not in ("fixed", "sha3"):
        raise ValueError("challenge_mode must be fixed or sha3")
    if layout not in ("element", "limb"):
        raise ValueError("layout must be element or limb")
    if point_mode not in ("generic", "specialized"):
        raise ValueError("point_mode must be generic or specialized")
    field = get_field(bit_width)
    spec = get_spec(workload)
    comp = Composition.from_spec(spec)
    ctx = ctx or default_context()
    n = 1 << rounds
    in_seed = seed + rounds if input_mode == "affine" else seed
    if layout == "limb":
        # the reference's limb-major tables ([table][limb][entry], field_sumcheck.cu:165-190):
        # the same generated inputs handed over limb-major and transposed on the device
        with Proof(ctx, bit_width, comp, rounds, input=input_mode, seed=in_seed) as gen:
            limb_major = np.ascontiguousarray(gen.tables().transpose(0, 2, 1))
        source = {"input": "host-mont-limb", "data": limb_major}
    else:
        source = {"input": input_mode, "seed": in_seed}
    fused = timing_mode == "round-sum"
    with Proof(ctx, bit_width, comp, rounds, **source) as proof:
        if fused:
            proof.set_profiling(True)
        first_ms = run_timed(proof, challenge_mode, timer)
        _p, _c, checksum = proof.result()
        for _ in range(warmups):
            run_timed(proof, challenge_mode, timer)
            checksum += proof.result()[2]
        samples, evals, challenges = [], [], []
        for _ in range(repeats):
            samples.append(run_timed(proof, challenge_mode, timer))
            checksum += proof.result()[2]
            if fused:
                e, c = phase_ms(proof)
                evals.append(e)
                challenges.append(c)
        polys, chal, _cs = proof.result()
    median_ms = statistics.median(samples) if samples else first_ms
    sorted_s = sorted(samples) if samples else [first_ms]
    p99 = sorted_s[min(len(sorted_s) - 1, int(round(0.99 * (len(sorted_s) - 1))))]
    algo = algorithmic_bytes(spec.num_vars, field.limbs, rounds)
    if fused and evals:
        eval_col, chal_col = f"{statistics.median(evals):.3f}", f"{statistics.median(challenges):.6f}"
        total_col = f"{statistics.median(samples):.3f}"
        fold_col = ""  # no separate fold launch: each round's fold is fused into the next pass
    else:  # whole-run timing: the reference writes the phase columns as zero
        eval_col, chal_col, fold_col, total_col = "0.000", "0.000000", "0.000", f"{median_ms:.3f}"
    gbps = algo / (median_ms * 1e6) if median_ms > 0 else float("inf")
    row = {
        "workload": spec.name,
        "polynomial": spec.polynomial,
        "backend": f"hip-sumcheck-prime-field-montgomery-{layout}-layout-{point_mode}-points",
        "field": field.name,
        "arithmetic": "montgomery",
        "layout": layout,
        "point_mode": point_mode,
        "bit_width": bit_width,
        "limbs": field.limbs,
        "modulus_hex": hex(field.modulus),
        "rounds": rounds,
        "N": n,
        "vars": spec.num_vars,
        "degree": spec.degree,
        "terms": spec.terms,
        "warmups": warmups,
        "repeats": repeats,
        "timing_mode": timing_mode,
        "challenge_mode": challenge_mode,
        "first_ms": f"{first_ms:.3f}",
        "median_ms": f"{median_ms:.3f}",
        "eval_median_ms": eval_col,
        "challenge_median_ms": chal_col,
        "fold_median_ms": fold_col,
        "total_median_ms": total_col,
        "input_mib": f"{spec.num_vars * n * field.limbs * 4 / (1024 ** 2):.6f}",
        "checksum": checksum,
        "status": "ok",
        "proofs_per_s": f"{1000.0 / median_ms:.3f}" if median_ms > 0 else "inf",
        "p50_ms": f"{median_ms:.3f}",
        "p99_ms": f"{p99:.3f}",
        "algo_bytes": algo,
        "hbm_gbps": f"{gbps:.1f}",
        "roofline_frac": f"{gbps / HBM_PEAK_GBPS:.4f}",
        "round_polys": polys,
        "challenges": chal,
    }
    return row
