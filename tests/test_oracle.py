"""CPU tests that pin the oracle (oracle/sumcheck_ref.py) before it is trusted as a checker.

Each check is chosen so that a plausible mistake (a dropped term, a wrong table numbering, a
wrong transcript byte, a swapped even/odd, an off-by-one challenge index) fails it:

* the reference's own recorded checksums (384 registry goldens, all 32 workloads x 4 widths x
  {r3 fixed, r10 fixed, r10 sha3});
* brute force from the definition of g_j via the multilinear extension on tiny sizes;
* the protocol invariants g_0(0)+g_0(1) = claim and g_j(0)+g_j(1) = g_{j-1}(r_{j-1});
* closed forms for Montgomery arithmetic, the published SplitMix64 sequence.
"""

from __future__ import annotations

import itertools
import json
import random
from pathlib import Path

import pytest

from oracle import sumcheck_ref as R
from zkduel_hip.workloads import SPEC_NAMES, get_spec

GOLDEN = Path(__file__).resolve().parent / "golden"
REGISTRY = json.loads((GOLDEN / "registry_checksums.json").read_text())


def spec_terms(name):
    s = get_spec(name)
    return s.num_vars, [(c, list(s.term_vars[o : o + n])) for c, n, o in zip(s.coeffs, s.term_lens, s.term_offsets)]


def oracle_checksum(workload, bit_width, rounds, challenge_mode, seed=1):
    """The registry validator's procedure (validate_vs_registry.py:94-148) on the oracle."""
    p, L = R.field(bit_width)
    T, terms = spec_terms(workload)
    tables = R.affine_tables(T, rounds, seed + rounds, p)
    ch = R.sha3_challenger(p, L) if challenge_mode == "sha3" else R.fixed_challenger(1)
    _claim, rounds_g, _rs = R.prove(tables, terms, p, rounds, ch)
    return R.checksum(rounds_g, p, L)


def test_registry_golden_count():
    assert len(REGISTRY) == 384
    assert {g["workload"] for g in REGISTRY} == set(SPEC_NAMES)


@pytest.mark.parametrize("bit_width", [32, 64, 128, 256])
def test_oracle_matches_reference_registry_goldens(bit_width):
    for g in REGISTRY:
        if g["bit_width"] != bit_width:
            continue
        got = 2 * oracle_checksum(g["workload"], bit_width, g["rounds"], g["challenge_mode"])
        assert got == g["checksum_x2"], g


def _mle(table, z, p):
    """Multilinear extension at z; variable i is bit i of the index (pairs (2i, 2i+1) fold first)."""
    total = 0
    for b, val in enumerate(table):
        w = 1
        for i, zi in enumerate(z):
            w = w * (zi if (b >> i) & 1 else (1 - zi)) % p
        total += val * w
    return total % p


@pytest.mark.parametrize("bit_width", [32, 64, 128, 256])
@pytest.mark.parametrize("num_vars", [1, 2, 3])
def test_prove_matches_brute_force_definition(bit_width, num_vars):
    p, L = R.field(bit_width)
    rng = random.Random(1000 * bit_width + num_vars)
    terms = [(rng.randrange(p), [0, 1, 1]), (p - 3, [2]), (5, [0, 2, 1, 0])]
    tables = [[rng.randrange(p) for _ in range(1 << num_vars)] for _ in range(3)]
    chal = [rng.randrange(p) for _ in range(num_vars)]
    claim, rounds, rs = R.prove(tables, terms, p, num_vars, lambda j, g: chal[j])
    d = 4
    assert rs == chal[: num_vars - 1]
    for j in range(num_vars):
        for t in range(d + 1):
            want = 0
            for x in itertools.product((0, 1), repeat=num_vars - j - 1):
                z = list(chal[:j]) + [t] + list(x)
                vals = [_mle(tb, z, p) for tb in tables]
                want += R._compose(terms, vals, p)
            assert rounds[j][t] == want % p, (j, t)
    assert claim == sum(R._compose(terms, [tb[i] for tb in tables], p) for i in range(1 << num_vars)) % p


@pytest.mark.parametrize("bit_width", [32, 64, 128, 256])
def test_protocol_invariants(bit_width):
    p, L = R.field(bit_width)
    T, terms = spec_terms("zk_vanilla_permcheck_hp")
    tables = R.uniform_tables(T, 6, seed=bit_width, p=p, limbs=L)
    claim, rounds, rs = R.prove(tables, terms, p, 6, R.sha3_challenger(p, L))
    assert (rounds[0][0] + rounds[0][1]) % p == claim
    for j in range(1, 6):
        assert (rounds[j][0] + rounds[j][1]) % p == R.lagrange_eval(rounds[j - 1], rs[j - 1], p)


def test_sumcheck_reference_contract():
    """reference.py:44-110's signature: tables by name, expression terms, challenge list."""
    p, _L = R.field(256)
    rng = random.Random(11)
    expr = [["a", "b", "c"], (-7, ["d", "e"])]
    tables = {v: [rng.randrange(p) for _ in range(16)] for v in "abcde"}
    chal = [rng.randrange(p) for _ in range(4)]
    claim, evals = R.sumcheck_reference(tables, q=p, expression=expr, challenges=chal, num_rounds=4)
    assert len(evals) == 4 and all(len(g) == 4 for g in evals)
    assert (evals[0][0] + evals[0][1]) % p == claim
    assert R.expression_vars(expr) == ["a", "b", "c", "d", "e"]
    assert R.expression_degree(expr) == 3


def test_lagrange_eval_closed_form():
    p = R.field(64)[0]
    coeffs = [5, p - 2, 7, 11]
    f = lambda x: sum(c * pow(x, k, p) for k, c in enumerate(coeffs)) % p
    vals = [f(t) for t in range(4)]
    for x in (0, 3, 17, p - 1, 123456789):
        assert R.lagrange_eval(vals, x, p) == f(x)


@pytest.mark.parametrize("bit_width", [32, 64, 128, 256])
def test_kernel_mont_mul_closed_form(bit_width):
    """The word routine equals a*b*R^-1 mod p on edges and stressed operands
    (src/zkduel/fields.py:154-187; the four fields are inside the bound of :118-121)."""
    p, L = R.field(bit_width)
    assert R.cios_bound_holds(p, L)
    rinv = pow(1 << (32 * L), -1, p)
    edge = [0, 1, 2, p - 2, p - 1, (1 << (32 * L)) % p]
    rng = random.Random(bit_width)
    pairs = [(a, b) for a in edge for b in edge]
    pairs += [(p - 1 - rng.randrange(256), rng.randrange(p)) for _ in range(200)]
    pairs += [(rng.randrange(p), rng.randrange(p)) for _ in range(200)]
    for a, b in pairs:
        assert R.kernel_mont_mul(a, b, p, L) == a * b * rinv % p


def test_kernel_mont_mul_unsafe_moduli_are_caught():
    """Moduli above the bound make the truncated routine wrong (tests/test_field_guard.py:94-115)."""
    for p, L in (((1 << 64) - 59, 2), ((1 << 128) - 159, 4)):
        assert not R.cios_bound_holds(p, L)
        rinv = pow(1 << (32 * L), -1, p)
        a = b = p - 1
        rng = random.Random(7)
        found = False
        for _ in range(2000):
            a, b = p - 1 - rng.randrange(1 << 8), p - 1 - rng.randrange(1 << 8)
            if R.kernel_mont_mul(a, b, p, L) != a * b * rinv % p:
                found = True
                break
        assert found


def test_field_constants_closed_form():
    """Literal vectors of cuda/src/field_sumcheck.cu:50-121 (nprime, R^2, k*R)."""
    p, L = R.field(256)
    assert R.nprime(p) == 0xFFFFFFFF
    r2 = (1 << 512) % p
    assert R.to_limbs(r2, L)[0] == 0xF3F29C6D and R.to_limbs(r2, L)[7] == 0x0748D9D9
    assert R.to_limbs(R.mont(3, p, L), L)[:3] == [0xFFFFFFFA, 0x00000005, 0x0009D806]
    assert R.nprime(R.field(32)[0]) == 0xCCCCCCCD
    assert R.nprime(R.field(128)[0]) == 0x00000001
    assert R.to_limbs(R.mont(7, *R.field(64)), 2) == [0xFFFFFFF9, 0x00000006]


def test_splitmix64_published_sequence():
    """SplitMix64 seeded with 0 (Steele, Lea, Flood 2014): the first three outputs."""
    s, out = 0, []
    for _ in range(3):
        s = (s + R.GAMMA) & R.M64
        out.append(R.mix64(s))
    assert out == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


@pytest.mark.parametrize("bit_width", [32, 64, 128, 256])
def test_uniform_generator_range_and_rejection(bit_width):
    p, L = R.field(bit_width)
    key = R.stream_key(5, 3, 2)
    vals = [R.uniform_element(key, i, p, L) for i in range(4000)]
    assert all(0 <= v < p for v in vals)
    assert len(set(vals)) == len(vals)
    mask = (1 << p.bit_length()) - 1
    # the first attempt is rejected exactly when the masked draw is >= p
    rejected = 0
    for i in range(4000):
        words = []
        for w in range((L + 1) // 2):
            d = R._draw(key, i * 128 + w)
            words += [d & R.M32, d >> 32]
        if R.from_limbs(words[:L]) & mask >= p:
            rejected += 1
            assert vals[i] != R.from_limbs(words[:L]) & mask
    if bit_width == 256:
        assert 250 < rejected < 500  # 1 - p / 2^255 = 0.094
    sel = [R.selector_element(key, i) for i in range(1000)]
    assert set(sel) == {0, 1}


def test_transcript_layout_is_little_endian_words():
    """One hand-built SHA3 message: tag, u32le round, u32le Montgomery limbs (triton_sweep.py:474-484)."""
    import hashlib

    p, L = R.field(64)
    g = [3, 5, 7, 11]
    msg = b"zkduel:field-sweep:v1" + (2).to_bytes(4, "little")
    for v in g:
        m = v * (1 << 64) % p
        msg += (m & R.M32).to_bytes(4, "little") + (m >> 32).to_bytes(4, "little")
    want = int.from_bytes(hashlib.sha3_256(msg).digest(), "little") % p
    assert R.transcript_challenge(2, g, p, L) == want
