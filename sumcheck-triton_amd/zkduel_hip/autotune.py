"""Schedule autotuner: picks a proof's launch schedule, and a batch's tile size, by timing the
candidates on the device for the actual (width, composition, size).

Knobs, none of which changes any result bit:

* ``max_pass_blocks`` — grid cap of the per-round pass kernel (the kernel loops over pairs);
* ``tail_pairs`` — rounds with at most this many pairs are fused into the single-workgroup tail
  kernel (the "fusion depth": how many final rounds run in one launch);
* ``items`` (batches) — pairs per thread in a batch tile.

Results are cached per (bit_width, num_tables, degree, num_vars) for the process, and, when
``path`` is given, in a JSON file, so a serving process can tune once per shape class.
"""

from __future__ import annotations

import json
import os
from pathlib import Path
from typing import Iterable

from .prover import Batch, Proof

PASS_BLOCKS = (512, 1024, 2048, 4096)
TAIL_PAIRS = (64, 256, 512, 2048)

_cache: dict[str, dict] = {}


def _key(proof: Proof) -> str:
    return f"{proof.field.bit_width}/{proof.comp.num_tables}/{proof.comp.degree}/{proof.num_vars}"


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


def tune_proof(proof: Proof, challenge_mode: str = "sha3", repeats: int = 3,
               pass_blocks: Iterable[int] = PASS_BLOCKS, tail_pairs: Iterable[int] = TAIL_PAIRS,
               path: str | os.PathLike | None = None) -> dict:
    """Sets the fastest (max_pass_blocks, tail_pairs) on ``proof`` and returns
    {"max_pass_blocks", "tail_pairs", "ms", "candidates": {"blocks/tail": ms}}. A cached
    choice for the same shape is applied without timing."""
    p = Path(path) if path is not None else None
    _load(p)
    key = _key(proof)
    if key in _cache:
        choice = _cache[key]
        proof.set_schedule(choice["max_pass_blocks"], choice["tail_pairs"], False)
        return choice
    timings: dict[str, float] = {}
    best = None
    # the two knobs are nearly independent: sweep the grid cap at the default tail, then the
    # tail at the best grid cap
    for blocks in pass_blocks:
        proof.set_schedule(blocks, 0, False)
        timings[f"{blocks}/default"] = _time(proof, challenge_mode, repeats)
    b_best = min(pass_blocks, key=lambda b: timings[f"{b}/default"])
    for tail in tail_pairs:
        proof.set_schedule(b_best, tail, False)
        ms = _time(proof, challenge_mode, repeats)
        timings[f"{b_best}/{tail}"] = ms
        if best is None or ms < best[2]:
            best = (b_best, tail, ms)
    choice = {"max_pass_blocks": best[0], "tail_pairs": best[1], "ms": best[2], "candidates": timings}
    proof.set_schedule(best[0], best[1], False)
    _cache[key] = choice
    _store(p)
    return choice


def tune_batch(batch: Batch, challenge_mode: str = "sha3", items: Iterable[int] = (1, 2, 4, 8),
               repeats: int = 1) -> dict[int, float]:
    """Tile size of a batch (pairs per thread), chosen by timing the actual mix."""
    return batch.autotune(tuple(items), challenge_mode, repeats)
