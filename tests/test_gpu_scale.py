"""GPU parity at the benchmark sizes and on the exact benchmark instances.

* Every checksum the reference recorded at scale (rounds 14..28, affine inputs, r = 1, 13 runs
  summed; experiments/h100/sweep_fixed/*.csv), including the rounds-28 rows whose tables hold
  more than 2^31 32-bit words (the size where the reference's int32 indexing broke,
  experiments/h100/FINDINGS.md:156-170).
* The benchmark instances with uniform tables and SHA3 challenges against the oracle's C twin:
  the Goldilocks a*b*c num_vars 16 proof, a*b*c + d*e at num_vars 20-22 at every width, and the
  BLS12-381 constraint workload, round polynomials, challenges and final point bit for bit.
* At the full benchmark sizes (num_vars 26 / 24), the SumCheck verifier with independent ends:
  the claim g_0(0) + g_0(1) and every table's multilinear extension at the challenge point are
  computed by the C twin from the same generated tables, every challenge is re-derived from its
  round polynomial, and every round is chained to the next by interpolation. A wrong round
  polynomial passes this with probability at most degree / p per round.
* The batched mix of 256 uneven instances against the same instances proved one at a time.
* Single-factor terms with coefficient -1 in degree-6 and degree-7 compositions (their largest
  interpolated value in the 255-bit working form) on the per-pair, point-parallel and batched
  kernels, against the Python oracle.
"""

from __future__ import annotations

import json
import random
import sys
from pathlib import Path

import numpy as np
import pytest

from oracle import cbind as C
from oracle import sumcheck_ref as R
from zkduel_hip import Composition, Context, Proof, get_field, get_spec, sumcheck
from zkduel_hip.prover import Batch
from zkduel_hip.synthetic import batch_instances, constraint_composition

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
SWEEP = json.loads((GOLDEN / "sweep_checksums.json").read_text())
WIDTHS = [32, 64, 128, 256]
SEED = 2026


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def _covered_elsewhere(g):
    # test_gpu_parity.py runs rounds 14 of every workload and a*b*c + d*e up to rounds 26
    return g["rounds"] == 14 or (g["workload"] == "poly_abc_plus_de" and g["rounds"] <= 26)


@pytest.mark.parametrize("bw", WIDTHS)
def test_every_recorded_sweep_checksum(ctx, bw):
    """All remaining recorded rows: every workload at rounds 16..28 (rows with 9 or more tables
    at rounds 28 are not recorded: the reference's own runs overflowed there)."""
    cases = sorted((g for g in SWEEP if g["bit_width"] == bw and not _covered_elsewhere(g)),
                   key=lambda g: (g["rounds"], g["workload"]))
    assert cases
    big = 0
    for g in cases:
        spec = get_spec(g["workload"])
        words = spec.num_vars * (1 << g["rounds"]) * (bw // 32)
        big += words >= 1 << 31
        say("sweep", g["workload"], bw, g["rounds"], f"{words / 2**31:.2f} x 2^31 words")
        comp = Composition.from_spec(spec)
        with Proof(ctx, bw, comp, g["rounds"], input="affine", seed=1 + g["rounds"]) as pr:
            pr.run("fixed")
            pr.wait()
            _p, _c, cs = pr.result()
        assert cs * g["runs"] == g["checksum_sum"], g
    if bw >= 64:
        assert big > 0  # some proofs index past 2^31 words


def _terms(comp):
    return [(c, list(comp.term_vars[o : o + n])) for c, n, o in zip(comp.coeffs, comp.term_lens, comp.term_offsets)]


def _against_c_twin(ctx, bw, comp, nv, selector_mask=0, seed=SEED):
    with Proof(ctx, bw, comp, nv, input="uniform", seed=seed, selector_mask=selector_mask) as pr:
        pr.run("sha3")
        pr.wait()
        polys, chal, _cs = pr.result()
        rf, finals = pr.final()
    want_polys, want_r, want_final = C.prove(bw, comp.num_tables, _terms(comp), nv, input_kind="uniform", seed=seed,
                                             selector_mask=selector_mask, challenge_mode="sha3")
    assert np.array_equal(polys, want_polys), "round polynomials"
    assert np.array_equal(chal, want_r), "challenges"
    # the C twin stops at the last pair of every table; fold it with the final challenge
    p, L = R.field(bw)
    r = R.unmont(R.from_limbs(rf), p, L)
    for v in range(comp.num_tables):
        lo, hi = (R.unmont(R.from_limbs(want_final[v][k]), p, L) for k in (0, 1))
        assert R.unmont(R.from_limbs(finals[v]), p, L) == (lo + r * (hi - lo)) % p, v


def test_goldilocks_abc_nv16_benchmark_instance(ctx):
    """The 16-round Goldilocks a*b*c proof exactly as the benchmark runs it (3 uniform
    tables of 2^16, SHA3), against the Python oracle and the C twin."""
    f = get_field(64)
    comp = Composition.from_spec(get_spec("poly_abc"))
    with Proof(ctx, 64, comp, 16, input="uniform", seed=SEED) as pr:
        pr.run("sha3")
        pr.wait()
        g, r, _cs = pr.canonical_result()
    tables = R.uniform_tables(3, 16, SEED, f.modulus, f.limbs)
    claim, want, want_r = R.prove(tables, comp.terms(), f.modulus, 16, R.sha3_challenger(f.modulus, f.limbs))
    assert g == want and r == want_r
    assert (g[0][0] + g[0][1]) % f.modulus == claim
    _against_c_twin(ctx, 64, comp, 16)


@pytest.mark.parametrize("bw,nv", [(32, 22), (64, 22), (128, 20), (256, 20)])
def test_abc_plus_de_uniform_sha3_against_c_twin(ctx, bw, nv):
    say("c-twin a*b*c+d*e", bw, nv)
    _against_c_twin(ctx, bw, Composition.from_spec(get_spec("poly_abc_plus_de")), nv)


@pytest.mark.parametrize("nv", [14, 16])
def test_bls_constraint_workload_against_c_twin(ctx, nv):
    """The benchmark's BLS12-381 constraint composition (10 uniform + 2 selector tables, 12
    monomials of degree 3-6 with random field coefficients)."""
    comp, mask = constraint_composition(256)
    say("c-twin constraint", nv)
    _against_c_twin(ctx, 256, comp, nv, selector_mask=mask)


@pytest.mark.parametrize("bw,nv", [(32, 26), (64, 26), (128, 24), (256, 24)])
def test_full_size_verifier_with_independent_claim_and_final_point(ctx, bw, nv):
    f = get_field(bw)
    p, L = f.modulus, f.limbs
    comp = Composition.from_spec(get_spec("poly_abc_plus_de"))
    say("full size", bw, nv)
    with Proof(ctx, bw, comp, nv, input="uniform", seed=SEED) as pr:
        pr.run("sha3")
        pr.wait()
        g, r, _cs = pr.canonical_result()
        polys, chal, _ = pr.result()
        rf, finals = pr.final()
    r_final = R.unmont(R.from_limbs(rf), p, L)
    point = np.concatenate([chal, rf.reshape(1, L)]).astype(np.uint32)
    say("c-twin claim and final point", bw, nv)
    claim_m, final_m = C.claim_and_final(bw, comp.num_tables, _terms(comp), nv, point, seed=SEED)
    claim = R.unmont(R.from_limbs(claim_m), p, L)
    assert (g[0][0] + g[0][1]) % p == claim
    for j in range(nv - 1):
        assert r[j] == R.transcript_challenge(j, g[j], p, L), j
    assert r_final == R.transcript_challenge(nv - 1, g[nv - 1], p, L)
    for j in range(1, nv):
        assert (g[j][0] + g[j][1]) % p == R.lagrange_eval(g[j - 1], r[j - 1], p), j
    assert np.array_equal(finals, final_m), "final values"
    vals = [R.unmont(R.from_limbs(x), p, L) for x in finals]
    a, b, c, d, e = vals
    assert R.lagrange_eval(g[nv - 1], r_final, p) == (a * b * c + d * e) % p


@pytest.mark.parametrize("bw", [64, 256])
def test_batched_mix_of_256_instances_equals_plain(ctx, bw):
    """The benchmark's batched mix (256 instances, num_vars 10..22, random product sums over
    2..8 tables), every instance's round polynomials and challenges equal to its plain proof."""
    insts = batch_instances(256, seed=4)
    with Batch(ctx, bw, insts) as b:
        b.run("sha3")
        b.wait()
        batched = [b.result(i) for i in range(len(insts))]
    for i, (comp, nv, inp) in enumerate(insts):
        if i % 32 == 0:
            say("plain", bw, i, nv)
        with Proof(ctx, bw, comp, nv, input="uniform", seed=inp["seed"], instance=inp["instance"]) as pr:
            pr.run("sha3")
            pr.wait()
            want = pr.result()
        assert np.array_equal(batched[i][0], want[0]) and np.array_equal(batched[i][1], want[1]), i


MINUS_TERMS = {
    # a b c d e f - g (degree 6) and a b c d e f h - g (degree 7), and the same over 10 tables
    "deg6": (7, [(1, [0, 1, 2, 3, 4, 5]), (-1, [6])]),
    "deg7": (8, [(1, [0, 1, 2, 3, 4, 5, 7]), (-1, [6])]),
    "deg7_wide": (10, [(1, [0, 1, 2, 3, 4, 5, 7]), (-1, [6]), (-1, [8]), (5, [9, 9])]),
}


@pytest.mark.parametrize("bw", WIDTHS)
@pytest.mark.parametrize("name", sorted(MINUS_TERMS))
def test_lone_factor_with_minus_one_coefficient(ctx, bw, name):
    f = get_field(bw)
    p, L = f.modulus, f.limbs
    T, terms = MINUS_TERMS[name]
    comp = Composition.from_terms(T, terms)
    nv = 8
    tables = R.uniform_tables(T, nv, 21, p, L)
    _claim, want, want_r = R.prove(tables, comp.terms(), p, nv, R.sha3_challenger(p, L))
    for wide in (0, 1 << 40, -1):  # per-pair kernels, point-parallel rounds, the default
        with Proof(ctx, bw, comp, nv, input="uniform", seed=21) as pr:
            pr.set_schedule(0, 1, False)  # only the last round in the tail kernel
            pr.set_point_parallel(wide)
            pr.run("sha3")
            pr.wait()
            g, r, _cs = pr.canonical_result()
        say("minus-one term", name, bw, wide)
        assert g == want and r == want_r, wide
    # a batch's tail takes the rounds of at most 512 pairs (64 at 255 bits): instance 0 is
    # large enough to run its first rounds in the stage kernels
    insts = [(comp, 11, {"kind": "uniform", "seed": 21, "instance": 0}),
             (comp, 6, {"kind": "uniform", "seed": 21, "instance": 1})]
    wants = []
    for nvi, inst in ((11, 0), (6, 1)):
        tab = R.uniform_tables(T, nvi, 21, p, L, instance=inst)
        wants.append(R.prove(tab, comp.terms(), p, nvi, R.sha3_challenger(p, L))[1])
    for wide in (0, 1 << 40):
        with Batch(ctx, bw, insts) as b:
            b.set_point_parallel(wide)
            b.run("sha3")
            b.wait()
            for i, w in enumerate(wants):
                polys, _chal, _cs = b.result(i)
                got = [[f.montgomery_decode(v) for v in f.limbs_to_ints(polys[j])] for j in range(len(w))]
                assert got == w, (wide, i)


def test_package_api_with_numpy_integer_coefficients():
    """Expression coefficients that int() converts (numpy integers) are accepted and mean the
    same as Python ints, as in the reference's _expr.py:11-24."""
    q = get_field(256).modulus
    expr = [(np.int64(3), ["a", "b"]), (np.int32(-2), ["c"]), ["a", "c", "c"]]
    rng = random.Random(7)
    tables = {v: [rng.randrange(q) for _ in range(1 << 5)] for v in ("a", "b", "c")}
    chal = [rng.randrange(q) for _ in range(5)]
    want = R.sumcheck_reference(tables, q=q, expression=[(3, ["a", "b"]), (-2, ["c"]), ["a", "c", "c"]],
                                challenges=chal, num_rounds=5)
    assert R.sumcheck_reference(tables, q=q, expression=expr, challenges=chal, num_rounds=5) == want
    claim, evals = sumcheck(tables, q=q, expression=expr, challenges=chal, num_rounds=5, bit_width=256, backend="hip")
    assert claim == want[0] and evals.tolist() == want[1]
