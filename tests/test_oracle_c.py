"""CPU tests pinning the oracle's C twin (oracle/oracle.c, used for num_vars 16..26):
against the Python oracle, hashlib, big-integer products, the reference's registry goldens and
a sample of its at-scale sweep checksums."""

from __future__ import annotations

import hashlib
import json
import random
from pathlib import Path

import numpy as np
import pytest

from oracle import cbind as C
from oracle import sumcheck_ref as R
from zkduel_hip.workloads import get_spec

GOLDEN = Path(__file__).resolve().parent / "golden"


def terms_of(name):
    s = get_spec(name)
    return s.num_vars, [(c, list(s.term_vars[o : o + n])) for c, n, o in zip(s.coeffs, s.term_lens, s.term_offsets)]


def checksum_of(polys):
    L = polys.shape[-1]
    return int(sum((i + 1) * int(polys[..., i].astype(np.uint64).sum()) for i in range(L)))


def test_c_sha3_matches_hashlib():
    rng = random.Random(1)
    for n in (0, 1, 57, 135, 136, 137, 271, 272, 1000):
        data = bytes(rng.randrange(256) for _ in range(n))
        assert C.sha3_256(data) == hashlib.sha3_256(data).digest(), n


@pytest.mark.parametrize("bw", [32, 64, 128, 256])
def test_c_montgomery_product(bw):
    p, L = R.field(bw)
    rng = random.Random(bw)
    a = [0, 1, p - 1, p - 2] + [rng.randrange(p) for _ in range(2000)]
    b = [p - 1, p - 1, 1, p - 2] + [rng.randrange(p) for _ in range(2000)]
    rinv = pow(1 << (32 * L), -1, p)
    assert C.mont_mul(bw, a, b) == [x * y * rinv % p for x, y in zip(a, b)]


@pytest.mark.parametrize("bw", [32, 64, 128, 256])
def test_c_twin_matches_python_oracle(bw):
    p, L = R.field(bw)
    for name in ("poly_abc_plus_de", "zk_complete_add_1", "zk_jellyfish_permcheck_hp"):
        T, terms = terms_of(name)
        for nv in (1, 3, 7):
            tables = R.uniform_tables(T, nv, 9, p, L, selectors=(1,))
            _claim, want, want_r = R.prove(tables, terms, p, nv, R.sha3_challenger(p, L))
            polys, rr, _fin = C.prove(bw, T, terms, nv, input_kind="uniform", seed=9, selector_mask=2,
                                      challenge_mode="sha3")
            got = [[R.unmont(R.from_limbs(polys[j][t]), p, L) for t in range(len(want[0]))] for j in range(nv)]
            assert got == want, (name, nv)
            assert [R.unmont(R.from_limbs(x), p, L) for x in rr] == want_r


def test_c_twin_reproduces_registry_goldens():
    gold = json.loads((GOLDEN / "registry_checksums.json").read_text())
    for g in gold:
        T, terms = terms_of(g["workload"])
        polys, _rr, _f = C.prove(g["bit_width"], T, terms, g["rounds"], input_kind="affine",
                                 seed=1 + g["rounds"], challenge_mode=g["challenge_mode"])
        assert 2 * checksum_of(polys) == g["checksum_x2"], g


@pytest.mark.parametrize("bw", [32, 64, 128, 256])
def test_c_twin_reproduces_sweep_checksum_at_scale(bw):
    """The reference's recorded a*b*c + d*e checksum at num_vars 18 (13 runs summed)."""
    sweep = json.loads((GOLDEN / "sweep_checksums.json").read_text())
    g = next(r for r in sweep if r["workload"] == "poly_abc_plus_de" and r["bit_width"] == bw and r["rounds"] == 18)
    T, terms = terms_of("poly_abc_plus_de")
    polys, _rr, _f = C.prove(bw, T, terms, 18, input_kind="affine", seed=19, challenge_mode="fixed")
    assert checksum_of(polys) * g["runs"] == g["checksum_sum"]


def _mle(values, point, p):
    """Multilinear extension by its definition: sum_i f_i prod_k (r_k if bit k of i else 1 - r_k)."""
    total = 0
    for i, x in enumerate(values):
        w = 1
        for k, r in enumerate(point):
            w = w * (r if (i >> k) & 1 else 1 - r) % p
        total += x * w
    return total % p


@pytest.mark.parametrize("bw", [32, 64, 128, 256])
def test_c_claim_and_final_point(bw):
    """oracle_claim_final (the full-size checks' claim and final point) against the Python
    oracle's claim and the multilinear extension by its definition, with a selector table."""
    p, L = R.field(bw)
    T, terms = terms_of("zk_complete_add_1")
    rng = random.Random(bw + 5)
    for nv in (1, 4, 7):
        point = [rng.randrange(p) for _ in range(nv)]
        tables = R.uniform_tables(T, nv, 11, p, L, selectors=(1,))
        claim, _g, _r = R.prove(tables, terms, p, nv, R.sha3_challenger(p, L))
        pm = [R.to_limbs(R.mont(x, p, L), L) for x in point]
        c, fin = C.claim_and_final(bw, T, terms, nv, pm, seed=11, selector_mask=2, threads=3)
        assert R.unmont(R.from_limbs(c), p, L) == claim, nv
        assert [R.unmont(R.from_limbs(x), p, L) for x in fin] == [_mle(t, point, p) for t in tables], nv
