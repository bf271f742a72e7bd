"""Package API, same contract as the reference's ``zkduel.sumcheck`` (src/zkduel/api.py:12-64).

``sumcheck(eval_tables, q=, expression=, challenges=, num_rounds=, bit_width=, backend=)``
returns ``(claim0, round_evals)`` with canonical Python ints, ``round_evals`` a
``(num_rounds, degree + 1)`` object array. The only backend is ``"hip"`` (the GPU library);
it accepts all four widths, where the reference's package backend accepts only 256.
"""

from __future__ import annotations

from typing import Iterable, Mapping, Sequence

import numpy as np

from . import _lib
from .expr import expression_vars
from .fields import get_field
from .prover import Composition, Proof, default_context

BACKENDS = ("hip",)


def available_backends() -> list[str]:
    """Usable backends: ``["hip"]`` when the library loads and a device is visible."""
    try:
        default_context()
    except Exception:
        return []
    return ["hip"]


def _tables_to_limbs(eval_tables, names: Sequence[str], num_rounds: int, field) -> np.ndarray:
    n = 1 << num_rounds
    out = np.empty((len(names), n, field.limbs), dtype=np.uint32)
    for v, name in enumerate(names):
        vals = eval_tables[name]
        if isinstance(vals, np.ndarray) and vals.dtype != object and field.limbs <= 2:
            arr = np.asarray(vals, dtype=np.uint64).reshape(-1)
            if arr.shape[0] != n:
                raise ValueError(f"Table for {name!r} has length {arr.shape[0]}, expected {n}")
            arr = arr % np.uint64(field.modulus) if field.modulus < (1 << 64) else arr
            out[v, :, 0] = (arr & np.uint64(0xFFFFFFFF)).astype(np.uint32)
            if field.limbs == 2:
                out[v, :, 1] = (arr >> np.uint64(32)).astype(np.uint32)
            continue
        vals = [int(x) for x in (vals.tolist() if isinstance(vals, np.ndarray) else vals)]
        if len(vals) != n:
            raise ValueError(f"Table for {name!r} has length {len(vals)}, expected {n}")
        out[v] = field.ints_to_limbs(vals)
    return out


def sumcheck(
    eval_tables: Mapping[str, Iterable[int] | np.ndarray],
    *,
    q: int,
    expression,
    challenges: Sequence[int],
    num_rounds: int,
    bit_width: int = 256,
    backend: str = "hip",
) -> tuple[int, np.ndarray]:
    """Run the SumCheck prover on the GPU with the caller's challenges (reference.py:44-110)."""
    if backend not in BACKENDS:
        raise ValueError(f"Unknown backend: {backend!r}; expected one of {list(BACKENDS)}")
    field = get_field(bit_width)
    if int(q) != field.modulus:
        raise ValueError(f"q={q} is not the {bit_width}-bit field modulus {field.modulus}")
    if num_rounds < 1:
        raise ValueError("num_rounds must be >= 1")
    comp, names = Composition.from_expression(expression)
    for name in expression_vars(expression):
        if name not in eval_tables:
            raise KeyError(f"missing table {name!r}")
    limbs = _tables_to_limbs(eval_tables, names, num_rounds, field)
    ctx = default_context()
    with Proof(ctx, bit_width, comp, num_rounds, input="host-canonical", data=limbs) as proof:
        proof.run("explicit", challenges)
        proof.wait()
        g, _r, _cs = proof.canonical_result()
    claim0 = (g[0][0] + g[0][1]) % field.modulus
    return int(claim0), np.array(g, dtype=object)


__all__ = ["available_backends", "sumcheck", "BACKENDS", "_lib"]
