"""Seeded synthetic workloads of the benchmarks (compositions, instance mixes, request streams).

All generators are pure functions of their seed (Python's ``random.Random``), so the GPU run,
the plain-vs-batched comparison and any CPU check see the same instances.
"""

from __future__ import annotations

import math
import random

from .fields import get_field
from .prover import Composition
from .workloads import get_spec


def abc_plus_de() -> Composition:
    """a*b*c + d*e over 5 tables (the at-scale plain workload)."""
    return Composition.from_spec(get_spec("poly_abc_plus_de"))


def constraint_composition(bit_width: int, seed: int = 3, tables: int = 10, selectors: int = 2,
                           monomials: int = 12) -> tuple[Composition, int]:
    """High-degree constraint-style composition (BLS12-381, num_vars 22 workload): ``tables`` uniform tables plus
    ``selectors`` selector tables, a sum of ``monomials`` products of degree 3..6 with seeded
    random field constants as coefficients. Every monomial carries one selector, as gate
    constraints do. Returns the composition and the selector mask (the last tables)."""
    rng = random.Random(seed)
    p = get_field(bit_width).modulus
    T = tables + selectors
    terms = []
    for k in range(monomials):
        deg = rng.randrange(3, 7)
        sel = tables + (k % selectors) if selectors else None
        factors = [rng.randrange(tables) for _ in range(deg - (1 if sel is not None else 0))]
        if sel is not None:
            factors.append(sel)
        terms.append((rng.randrange(1, p), factors))
    mask = sum(1 << v for v in range(tables, T))
    return Composition.from_terms(T, terms), mask


def random_product_sum(rng: random.Random) -> Composition:
    """A random sum of 1..4 products of degree 2..5 over 2..8 tables (the batched mix)."""
    T = rng.randrange(2, 9)
    K = rng.randrange(1, 5)
    terms = []
    for _ in range(K):
        deg = rng.randrange(2, 6)
        terms.append((1, [rng.randrange(T) for _ in range(deg)]))
    return Composition.from_terms(T, terms)


def batch_instances(n: int = 256, seed: int = 4, nv_lo: int = 10, nv_hi: int = 22) -> list[tuple]:
    """The batched mix: ``n`` instances, num_vars uniform in [nv_lo, nv_hi], random product sums,
    uniform tables keyed by instance number."""
    rng = random.Random(seed)
    out = []
    for i in range(n):
        comp = random_product_sum(rng)
        nv = rng.randrange(nv_lo, nv_hi + 1)
        out.append((comp, nv, {"kind": "uniform", "seed": seed, "instance": i}))
    return out


def serving_requests(n: int, rate_per_s: float, seed: int = 5, nv_lo: int = 12, nv_hi: int = 24):
    """The serving request stream: Poisson arrivals at ``rate_per_s``, num_vars log-uniform in
    [nv_lo, nv_hi] (uniform in log2(num_vars)), width 256/64/32 with probability 0.5/0.3/0.2,
    compositions as in the batched mix. Yields dicts with arrival time (s from start)."""
    rng = random.Random(seed)
    t = 0.0
    reqs = []
    for i in range(n):
        t += rng.expovariate(rate_per_s)
        u = rng.random()
        bw = 256 if u < 0.5 else (64 if u < 0.8 else 32)
        nv = int(round(math.exp(rng.uniform(math.log(nv_lo), math.log(nv_hi)))))
        reqs.append({"id": i, "arrival": t, "bit_width": bw, "num_vars": nv,
                     "composition": random_product_sum(rng), "seed": seed, "instance": i})
    return reqs
