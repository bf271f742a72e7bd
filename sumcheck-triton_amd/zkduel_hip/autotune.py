"""Schedule autotuner: picks a proof's launch schedule, and a batch's per-class schedules, by
timing the candidates on the device for the actual (width, composition, size).

Knobs, none of which changes any result bit:

* ``max_pass_blocks`` — grid cap of the per-round pass kernel (the kernel loops over pairs);
* ``tail_pairs`` — rounds with at most this many pairs are fused into the single-workgroup tail
  kernel (the "fusion depth": how many final rounds run in one launch);
* ``wide_pairs`` — rounds with at most this many pairs run as a fold launch plus the
  point-parallel evaluation instead of the per-pair pass kernel (-1: the library's default);
* ``items`` (batches and served requests) — pairs per thread in a stage tile, per (width,
  degree, size) class of instance (``tune_batch``; a Server applies a tuned class table to
  the requests it admits).

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

PASS_BLOCKS = (512, 768, 1024, 2048, 4096)
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


def instance_class(bit_width: int, comp, num_vars: int) -> str:
    """The (width, degree, size) class a batched or served instance is tuned by: degree 2, 3,
    4-5 or 6-7; num_vars up to 14, 15-18 or 19 and more; more than 8 tables or not."""
    d = comp.degree
    dc = "d2" if d <= 2 else ("d3" if d == 3 else ("d4-5" if d <= 5 else "d6-7"))
    sc = "n<=14" if num_vars <= 14 else ("n15-18" if num_vars <= 18 else "n>=19")
    tc = "t>8" if comp.num_tables > 8 else "t<=8"
    return f"{bit_width}/{dc}/{sc}/{tc}"


def _batch_ms(batch: Batch, mode: str, repeats: int) -> float:
    batch.run(mode)
    batch.wait()
    best = float("inf")
    for _ in range(max(1, repeats)):
        batch.run(mode)
        best = min(best, batch.wait())
    return best


BATCH_ITEMS = (1, 2, 4, 8, 16)
BATCH_WIDE = (-1, 0, 1 << 40)
BATCH_TAIL = (0, 16, 4096)


def tune_batch(batch: Batch, challenge_mode: str = "sha3", items: Iterable[int] = BATCH_ITEMS,
               wide_pairs: Iterable[int] = BATCH_WIDE, tail_pairs: Iterable[int] = BATCH_TAIL,
               repeats: int = 2, path: str | os.PathLike | None = None, margin: float = 0.01) -> dict:
    """Per-class schedules of a batch: for each (width, degree, size) class of its instances
    (instance_class), the tile (pairs per thread), the kernel choice (point-parallel threshold:
    -1 the library's rule, 0 per-pair, 2^40 every round point-parallel) and the fusion depth of
    its tail (pairs at or below which its last rounds run in the batch's one tail launch: 0 the
    default). Classes are tuned one at a time, largest work first, each candidate timed on the
    whole batch with the other classes at their current choice (the stage launches are
    shared, so a class's choice is judged by the batch's time); a candidate replaces the
    current choice when it is faster by more than ``margin``. Sets the choices on the batch
    and returns {"classes": {class: {"items", "max_pairs", "tail_pairs", "instances"}}, "ms",
    "default_ms", "candidates"}. Classes cached (per process, and in ``path`` when given) are
    applied first and not re-tuned."""
    p = Path(path) if path is not None else None
    _load(p)
    bw = batch.field.bit_width
    n = len(batch)
    cls_of = [instance_class(bw, batch.comps[i], batch.num_vars[i]) for i in range(n)]
    members: dict[str, list[int]] = {}
    for i, c in enumerate(cls_of):
        members.setdefault(c, []).append(i)
    work = {c: sum((batch.comps[i].num_tables * (1 << batch.num_vars[i])) for i in idx) for c, idx in members.items()}
    choice = {c: {"items": 4, "max_pairs": -1, "tail_pairs": 0} for c in members}
    for c in members:
        if f"batch/{c}" in _cache:
            choice[c] = {k: _cache[f"batch/{c}"][k] for k in ("items", "max_pairs", "tail_pairs")}

    def apply() -> None:
        batch.set_schedules([choice[c]["items"] for c in cls_of], [choice[c]["max_pairs"] for c in cls_of],
                            [choice[c]["tail_pairs"] for c in cls_of])

    apply()
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < 0.1:  # warm the device up before the first timing
        batch.run(challenge_mode)
        batch.wait()
    default_ms = _batch_ms(batch, challenge_mode, repeats)
    best_ms = default_ms
    candidates: dict[str, float] = {}
    for c in sorted(members, key=lambda c: -work[c]):
        if f"batch/{c}" in _cache:
            continue
        for knob, values in (("max_pairs", tuple(wide_pairs)), ("items", tuple(items)),
                             ("tail_pairs", tuple(tail_pairs))):
            current = choice[c][knob]
            for v in values:
                if v == current:
                    continue
                choice[c][knob] = v
                apply()
                ms = _batch_ms(batch, challenge_mode, repeats)
                candidates[f"{c}:{knob}={v}"] = round(ms, 3)
                if ms < best_ms * (1.0 - margin):
                    best_ms, current = ms, v
                choice[c][knob] = current
            apply()
        _cache[f"batch/{c}"] = dict(choice[c])
    _store(p)
    apply()
    final_ms = _batch_ms(batch, challenge_mode, repeats)
    return {"classes": {c: dict(choice[c], instances=len(members[c])) for c in sorted(members)},
            "ms": final_ms, "default_ms": default_ms, "candidates": candidates}
