"""Pure-Python SumCheck oracle — TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Every function restates the reference algorithm it cites; arithmetic is done on Python big
integers, so results are exact by construction. Nothing here is fast: it is the checker for
small sizes and the single-core CPU baseline, never the product path.

Parity pinning (what makes these functions trustworthy, exercised by ``tests/``):

* ``prove`` / ``sumcheck_reference``: the reference's 384 registry golden checksums
  (``experiments/h100/validation/*.jsonl``), brute force from the definition of g_j on
  num_vars <= 3, and the protocol invariants g_0(0)+g_0(1) = claim and
  g_j(0)+g_j(1) = g_{j-1}(r_{j-1}).
* ``transcript_challenge``: ``hashlib.sha3_256`` is the FIPS-202 function; the byte layout is
  pinned by the 128 SHA3-mode registry goldens.
* ``kernel_mont_mul``: closed form a*b*R^-1 mod p on edge and stressed operands, and the two
  unsafe moduli the reference rejects (``tests/test_field_guard.py`` of the reference).
* ``uniform_tables``: this project's own counter-based generator (SplitMix64, Steele-Lea-Flood
  2014); pinned by the published SplitMix64 output sequence and by a rejection-rate check.
"""

from __future__ import annotations

import hashlib
from typing import Callable, Iterable, Mapping, Sequence

M32 = 0xFFFFFFFF
M64 = 0xFFFFFFFFFFFFFFFF

# src/zkduel/fields.py:258-294 — bit width -> (modulus, limbs)
FIELDS: dict[int, tuple[int, int]] = {
    32: (0xFFFFFFFB, 1),
    64: (0xFFFFFFFF00000001, 2),
    128: ((1 << 127) - 1, 4),
    256: (int("73eda753299d7d483339d80809a1d80553bda402fffe5bfeffffffff00000001", 16), 8),
}

TRANSCRIPT_DOMAIN = b"zkduel:field-sweep:v1"  # cuda/src/field_sumcheck.cu:268


def field(bit_width: int) -> tuple[int, int]:
    return FIELDS[int(bit_width)]


def nprime(p: int) -> int:
    """-p^-1 mod 2^32 (src/zkduel/fields.py:50-52)."""
    return (-pow(p, -1, 1 << 32)) & M32


def to_limbs(x: int, limbs: int) -> list[int]:
    return [(x >> (32 * i)) & M32 for i in range(limbs)]


def from_limbs(words: Sequence[int]) -> int:
    return sum(int(w) << (32 * i) for i, w in enumerate(words))


def mont(x: int, p: int, limbs: int) -> int:
    """Montgomery image x*R mod p with R = 2^(32L) (src/zkduel/fields.py:251-252)."""
    return (x % p) * (1 << (32 * limbs)) % p


def unmont(x: int, p: int, limbs: int) -> int:
    return x * pow(1 << (32 * limbs), -1, p) % p


# ---------------------------------------------------------------------------------------------
# Expression rules: src/zkduel/_expr.py:27-67
# ---------------------------------------------------------------------------------------------

def _coefficient_term(term):
    """(int(c), names) when term is a two-item list or tuple whose first item is not a string
    and converts with int() (ints, bools, numpy integers, floats truncated) and whose second
    item is a non-string iterable of strings (_expr.py:11-24); else None."""
    if not isinstance(term, (list, tuple)) or len(term) != 2:
        return None
    c, names = term
    if isinstance(c, str) or isinstance(names, str):
        return None
    try:
        coeff = int(c)
        names = tuple(names)
    except (TypeError, ValueError):
        return None
    if not all(isinstance(v, str) for v in names):
        return None
    return coeff, names


def normalize_expression(expression) -> tuple[tuple[int, tuple[str, ...]], ...]:
    """Terms are ``[names]`` (coefficient 1) or ``(coeff, [names])`` (_expr.py:27-54)."""
    out = []
    for term in expression:
        weighted = _coefficient_term(term)
        if weighted is not None:
            out.append(weighted)
            continue
        if isinstance(term, str):
            raise TypeError("Expression terms must be sequences of variable names")
        factors = tuple(term)
        if not all(isinstance(v, str) for v in factors):
            raise TypeError("Expression terms must contain variable names")
        out.append((1, factors))
    if not out:
        raise ValueError("expression must contain at least one term")
    return tuple(out)


def expression_vars(expression) -> list[str]:
    """First-appearance order (_expr.py:57-63)."""
    seen: list[str] = []
    for _c, term in normalize_expression(expression):
        for v in term:
            if v not in seen:
                seen.append(v)
    return seen


def expression_degree(expression) -> int:
    """Longest term (_expr.py:66-67)."""
    return max(len(t) for _c, t in normalize_expression(expression))


# ---------------------------------------------------------------------------------------------
# The prover loop: src/zkduel/reference.py:44-110
# ---------------------------------------------------------------------------------------------

def _compose(terms: Sequence[tuple[int, Sequence[int]]], vals: Sequence[int], p: int) -> int:
    """One composition value (reference.py:17-26): sum_k c_k * prod_m vals[idx_km] mod p."""
    total = 0
    for c, idxs in terms:
        acc = c % p
        for i in idxs:
            acc = acc * vals[i] % p
        total += acc
    return total % p


def _sum_compose(terms, tables: Sequence[Sequence[int]], p: int) -> int:
    """sum over i of C(tables[.][i]) (reference.py:29-35 summed as at :88-99)."""
    n = len(tables[0])
    total = 0
    for i in range(n):
        total += _compose(terms, [t[i] for t in tables], p)
    return total % p


def prove(
    tables: Sequence[Sequence[int]],
    terms: Sequence[tuple[int, Sequence[int]]],
    p: int,
    num_rounds: int,
    challenge: Callable[[int, list[int]], int],
    degree: int | None = None,
) -> tuple[int, list[list[int]], list[int]]:
    """Round loop of reference.py:78-108 with the challenge supplied per round.

    ``terms`` are ``(coeff, table_indices)``; ``challenge(round, g_points)`` returns the
    canonical r for folding after that round (called only for rounds < num_rounds-1).
    Returns ``(claim0, round_evals[num_rounds][degree+1], challenges)`` in canonical form.
    """
    if degree is None:
        degree = max(len(idx) for _c, idx in terms)
    cur = [[int(x) % p for x in t] for t in tables]
    n = 1 << num_rounds
    for t in cur:
        if len(t) != n:
            raise ValueError(f"table has length {len(t)}, expected {n}")
    claim0 = _sum_compose(terms, cur, p)
    rounds: list[list[int]] = []
    rs: list[int] = []
    for j in range(num_rounds):
        evens = [t[0::2] for t in cur]
        odds = [t[1::2] for t in cur]
        diffs = [[(o - e) % p for o, e in zip(od, ev)] for od, ev in zip(odds, evens)]
        g = [_sum_compose(terms, evens, p), _sum_compose(terms, odds, p)]
        point = [list(o) for o in odds]
        for _t in range(2, degree + 1):  # incremental t >= 2 (reference.py:94-99)
            point = [[(c + d) % p for c, d in zip(pt, df)] for pt, df in zip(point, diffs)]
            g.append(_sum_compose(terms, point, p))
        rounds.append(g)
        if j < num_rounds - 1:
            r = challenge(j, g) % p
            rs.append(r)
            cur = [[(e + r * d) % p for e, d in zip(ev, df)] for ev, df in zip(evens, diffs)]
    return claim0, rounds, rs


def sumcheck_reference(
    eval_tables: Mapping[str, Iterable[int]],
    *,
    q: int,
    expression,
    challenges: Sequence[int],
    num_rounds: int,
) -> tuple[int, list[list[int]]]:
    """Same contract as reference.py:44-110: ``(claim0, round_evals)``, canonical ints."""
    names = expression_vars(expression)
    idx = {v: i for i, v in enumerate(names)}
    terms = [(c, [idx[v] for v in t]) for c, t in normalize_expression(expression)]
    tables = [list(eval_tables[v]) for v in names]
    claim0, rounds, _ = prove(
        tables, terms, q, num_rounds, lambda j, g: int(challenges[j]),
        degree=expression_degree(expression),
    )
    return claim0, rounds


# ---------------------------------------------------------------------------------------------
# Transcript, checksum, inputs
# ---------------------------------------------------------------------------------------------

def transcript_challenge(round_idx: int, g_points: Sequence[int], p: int, limbs: int) -> int:
    """Canonical r for ``round_idx`` (triton_sweep.py:474-484, field_sumcheck.cu:857-907).

    SHA3-256(domain || u32le(round) || u32le(every Montgomery limb of g(0..d))), digest read
    little-endian, reduced mod p. Stateless per round.
    """
    h = hashlib.sha3_256()
    h.update(TRANSCRIPT_DOMAIN)
    h.update(int(round_idx).to_bytes(4, "little"))
    for g in g_points:
        for w in to_limbs(mont(g, p, limbs), limbs):
            h.update(w.to_bytes(4, "little"))
    return int.from_bytes(h.digest(), "little") % p


def sha3_challenger(p: int, limbs: int) -> Callable[[int, list[int]], int]:
    return lambda j, g: transcript_challenge(j, g, p, limbs)


def fixed_challenger(value: int = 1) -> Callable[[int, list[int]], int]:
    """The reference's "fixed" mode folds with r = 1 (field_sumcheck.cu:1058)."""
    return lambda j, g: value


def checksum(rounds: Sequence[Sequence[int]], p: int, limbs: int) -> int:
    """sum over rounds, points, limbs of (i+1)*limb_i(Mont(g)) (field_sumcheck.cu:1043-1049)."""
    total = 0
    for g in rounds:
        for v in g:
            for i, w in enumerate(to_limbs(mont(v, p, limbs), limbs)):
                total += (i + 1) * w
    return total


def affine_tables(num_tables: int, num_vars: int, seed: int, p: int) -> list[list[int]]:
    """table[v][o] = seed + 7919(v+1) + 104729(o+1), reduced mod p (field_sumcheck.cu:123-128).

    The reference reduces only at 32 bits; at the other widths the value is already < p for
    every size it runs, so reducing unconditionally is the same function there.
    """
    n = 1 << num_vars
    return [[(seed + 7919 * (v + 1) + 104729 * (o + 1)) % p for o in range(n)] for v in range(num_tables)]


# ---------------------------------------------------------------------------------------------
# This project's counter-based input generator (restated identically in the C twin and in the
# device fill kernel). SplitMix64 finalizer: Steele, Lea, Flood, "Fast splittable
# pseudorandom number generators", OOPSLA 2014.
# ---------------------------------------------------------------------------------------------

GAMMA = 0x9E3779B97F4A7C15
MAX_ATTEMPTS = 32


def mix64(z: int) -> int:
    z &= M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def stream_key(seed: int, instance: int, table: int) -> int:
    k = mix64((seed & M64) ^ 0x5851F42D4C957F2D)
    return mix64(k ^ (((instance & M32) << 32) | (table & M32)))


def _draw(key: int, counter: int) -> int:
    return mix64((key + counter * GAMMA) & M64)


def uniform_element(key: int, index: int, p: int, limbs: int) -> int:
    """Rejection-sampled uniform value in [0, p): L words from draws, masked to bitlen(p)."""
    mask = (1 << p.bit_length()) - 1
    x = 0
    for attempt in range(MAX_ATTEMPTS):
        words = []
        for w in range((limbs + 1) // 2):
            d = _draw(key, index * 128 + attempt * 4 + w)
            words += [d & M32, d >> 32]
        x = from_limbs(words[:limbs]) & mask
        if x < p:
            return x
    return x % p


def selector_element(key: int, index: int) -> int:
    return _draw(key, index * 128) & 1


def uniform_tables(
    num_tables: int, num_vars: int, seed: int, p: int, limbs: int,
    instance: int = 0, selectors: Sequence[int] = (),
) -> list[list[int]]:
    n = 1 << num_vars
    out = []
    for v in range(num_tables):
        key = stream_key(seed, instance, v)
        if v in selectors:
            out.append([selector_element(key, i) for i in range(n)])
        else:
            out.append([uniform_element(key, i, p, limbs) for i in range(n)])
    return out


# ---------------------------------------------------------------------------------------------
# The kernels' word-level Montgomery routine: src/zkduel/fields.py:63-121
# ---------------------------------------------------------------------------------------------

def kernel_mont_mul(a: int, b: int, p: int, limbs: int) -> int:
    """Truncated CIOS with L+1 accumulator words and the top carry dropped (fields.py:63-115)."""
    L = limbs
    av, bv, mv = to_limbs(a, L), to_limbs(b, L), to_limbs(p, L)
    npr = nprime(p)
    t = [0] * (L + 1)
    for i in range(L):
        c = 0
        for j in range(L):
            uv = t[j] + av[j] * bv[i] + c
            t[j], c = uv & M32, uv >> 32
        t[L] = (t[L] + c) & M32  # bit 32 discarded, as in the kernels
        q = (t[0] * npr) & M32
        c = 0
        for j in range(L):
            uv = t[j] + q * mv[j] + c
            c = uv >> 32
            if j > 0:
                t[j - 1] = uv & M32
        last = t[L] + c
        t[L - 1], t[L] = last & M32, (last >> 32) & M32
    res = from_limbs(t[:L])
    if t[L] != 0 or res >= p:
        res = (res - p) & ((1 << (32 * L)) - 1)
    return res


def cios_bound_holds(p: int, limbs: int) -> bool:
    """(2p-1) + (p-1)(2^32-1) < 2^(32(L+1)) (fields.py:118-121)."""
    return (2 * p - 1) + (p - 1) * M32 < (1 << (32 * (limbs + 1)))


def lagrange_eval(values: Sequence[int], x: int, p: int) -> int:
    """Evaluate the degree-d polynomial with values g(0..d) at x (the verifier's step)."""
    d = len(values) - 1
    total = 0
    for i, yi in enumerate(values):
        num, den = 1, 1
        for k in range(d + 1):
            if k != i:
                num = num * (x - k) % p
                den = den * (i - k) % p
        total += yi * num * pow(den, -1, p)
    return total % p
