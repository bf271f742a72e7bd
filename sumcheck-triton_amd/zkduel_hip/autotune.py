"""Schedule autotuner: picks a proof's launch schedule, and a batch's tile size, by timing the
candidates on the device for the actual (width, composition, size).

Knobs, none of which changes any result bit:

* ``max_pass_blocks`` — grid cap of the per-round pass kernel (the kernel loops over pairs);
* ``tail_pairs`` — rounds with at most this many pairs are fused into the single-workgroup tail
  kernel (the "fusion depth": how many final rounds run in one launch);
* ``wide_pairs`` — rounds with at most this many pairs run as a fold launch plus the
  point-parallel evaluation instead of the per-pair pass kernel (-1: the library's default);
* ``items`` (batches) — pairs per thread in a batch tile.

The library's default schedule (0 / 0 / -1) is always a candidate and is kept unless the
best tuned schedule beats it by more than ``margin`` (2% by default): candidates closer than
that differ by timing noise, and a noisy pick would change the schedule from run to run.

Results are cached per (bit_width, composition, num_vars, compiled or interpreted kernels) for
the process, and, when ``path`` is given, in a JSON file, so a serving process can tune once
per shape class.
"""

from __future__ import annotations

import hashlib
import json
import os
import time
from pathlib import Path
from typing import Iterable

from .prover import Batch, Proof

PASS_BLOCKS = (512, 1024, 2048, 4096)
TAIL_PAIRS = (64, 256, 512, 2048)
WIDE_PAIRS = (-1, 0, 1 << 15, 1 << 17, 1 << 19, 1 << 21, 1 << 40)

_cache: dict[str, dict] = {}


def _key(proof: Proof) -> str:
    terms = hashlib.sha1(repr(proof.comp.terms()).encode()).hexdigest()[:12]
    return (f"{proof.field.bit_width}/{proof.comp.num_tables}/{proof.comp.degree}/{proof.num_vars}"
            f"/{terms}/{'compiled' if proof.compiled else 'interpreted'}")


def _load(path: Path | None) -> None:
    if path is not None and path.exists():
        try:
            _cache.update(json.loads(path.read_text()))
        except (OSError, ValueError):
            pass


def _store(path: Path | None) -> None:
    if path is None:
        return
    path.parent.mkdir(parents=True, exist_ok=True)
    tmp = path.with_suffix(".tmp")
    tmp.write_text(json.dumps(_cache, indent=1, sort_keys=True))
    os.replace(tmp, path)


def _time(proof: Proof, mode: str, repeats: int) -> float:
    proof.run(mode)
    proof.wait()
    best = float("inf")
    for _ in range(repeats):
        proof.run(mode)
        best = min(best, proof.wait())
    return best


def _apply(proof: Proof, choice: dict) -> None:
    proof.set_schedule(choice["max_pass_blocks"], choice["tail_pairs"], False)
    proof.set_point_parallel(choice.get("wide_pairs", -1))


def tune_proof(proof: Proof, challenge_mode: str = "sha3", repeats: int = 3,
               pass_blocks: Iterable[int] = PASS_BLOCKS, tail_pairs: Iterable[int] = TAIL_PAIRS,
               wide_pairs: Iterable[int] = WIDE_PAIRS,
               path: str | os.PathLike | None = None, margin: float = 0.02) -> dict:
    """Sets the fastest (max_pass_blocks, tail_pairs, wide_pairs) on ``proof`` and returns
    {"max_pass_blocks", "tail_pairs", "wide_pairs", "ms", "default_ms", "library_default",
    "candidates": {"blocks/tail/wide": ms}}; 0 / 0 / -1 with ``library_default`` True when the
    default schedule is kept. A cached choice for the same shape is applied without timing."""
    p = Path(path) if path is not None else None
    _load(p)
    key = _key(proof)
    if key in _cache:
        _apply(proof, _cache[key])
        return _cache[key]
    pass_blocks, tail_pairs, wide_pairs = tuple(pass_blocks), tuple(tail_pairs), tuple(wide_pairs)
    timings: dict[str, float] = {}

    def timed(blocks: int, tail: int, wide: int) -> float:
        name = f"{blocks}/{tail}/{wide}"
        if name not in timings:
            _apply(proof, {"max_pass_blocks": blocks, "tail_pairs": tail, "wide_pairs": wide})
            timings[name] = _time(proof, challenge_mode, repeats)
        return timings[name]

    # the device runs slow for the first few tens of milliseconds of work after idling; without
    # this the first candidate timed (the default) reads several percent slow
    _apply(proof, {"max_pass_blocks": 0, "tail_pairs": 0, "wide_pairs": -1})
    t0 = time.perf_counter()
    runs = 0
    while runs < 3 or time.perf_counter() - t0 < 0.1:
        proof.run(challenge_mode)
        proof.wait()
        runs += 1
    default_ms = timed(0, 0, -1)
    # the knobs are nearly independent: sweep the kernel choice at the default grid and tail,
    # then the grid cap, then the tail, each at the best of the ones before
    w_best = min(wide_pairs, key=lambda w: timed(0, 0, w))
    b_best = min(pass_blocks, key=lambda b: timed(b, 0, w_best))
    t_best = min(tail_pairs, key=lambda t: timed(b_best, t, w_best))
    best_ms = timings[f"{b_best}/{t_best}/{w_best}"]
    # timed again last, so the default is not compared on its first (earliest) timing alone
    _apply(proof, {"max_pass_blocks": 0, "tail_pairs": 0, "wide_pairs": -1})
    default_ms = min(default_ms, _time(proof, challenge_mode, repeats))
    timings["0/0/-1"] = default_ms
    keep_default = best_ms >= default_ms * (1.0 - margin)
    if keep_default:
        b_best, t_best, w_best, best_ms = 0, 0, -1, default_ms
    choice = {"max_pass_blocks": b_best, "tail_pairs": t_best, "wide_pairs": w_best,
              "ms": best_ms, "default_ms": default_ms, "library_default": keep_default,
              "candidates": timings}
    _apply(proof, choice)
    _cache[key] = choice
    _store(p)
    return choice


def tune_batch(batch: Batch, challenge_mode: str = "sha3", items: Iterable[int] = (1, 2, 4, 8, 16, 32),
               repeats: int = 1) -> dict[int, float]:
    """Tile size of a batch (pairs per thread), chosen by timing the actual mix."""
    return batch.autotune(tuple(items), challenge_mode, repeats)
