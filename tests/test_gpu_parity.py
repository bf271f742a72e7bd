"""GPU parity tests: the HIP library (through its C ABI) against the oracle and the reference's
recorded checksums. Bit-exact everywhere: this is integer arithmetic.

Sizes: the oracle runs full round polynomials up to num_vars 10; the reference's recorded
checksums cover num_vars 3..28 (affine inputs, fixed challenge); at the benchmark's full sizes
the size-independent protocol invariants are checked (g_0(0)+g_0(1) = claim is not available
without the oracle, so g_j(0)+g_j(1) = g_{j-1}(r_{j-1}) for every round, and each r_j is
re-derived from g_j by the oracle's transcript).
"""

from __future__ import annotations

import json
import random
import sys
from pathlib import Path

import numpy as np
import pytest

from oracle import sumcheck_ref as R
from zkduel_hip import Composition, Context, Proof, get_field, get_spec, sumcheck
from zkduel_hip.workloads import SPEC_NAMES

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
REGISTRY = json.loads((GOLDEN / "registry_checksums.json").read_text())
SWEEP = json.loads((GOLDEN / "sweep_checksums.json").read_text())
WIDTHS = [32, 64, 128, 256]


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def mont_ints(field, limbs_arr):
    return [field.montgomery_decode(v) for v in field.limbs_to_ints(limbs_arr)]


# ----------------------------------------------------------------------------- field layer

@pytest.mark.parametrize("bw", WIDTHS)
def test_field_ops_match_big_integers(ctx, bw):
    f = get_field(bw)
    p, L = f.modulus, f.limbs
    rng = random.Random(bw)
    edge = [0, 1, 2, p - 2, p - 1, f.r_mod, (1 << (32 * L - 1)) % p]
    a = edge * len(edge) + [rng.randrange(p) for _ in range(3000)] + [p - 1 - rng.randrange(256) for _ in range(500)]
    b = [y for y in edge for _ in edge] + [rng.randrange(p) for _ in range(3000)] + [p - 1 - rng.randrange(256) for _ in range(500)]
    A, B = f.ints_to_limbs(a), f.ints_to_limbs(b)
    rinv = f.r_inv
    say("field", bw, len(a))
    assert f.limbs_to_ints(ctx.field_op(bw, "mul", A, B)) == [x * y * rinv % p for x, y in zip(a, b)]
    assert f.limbs_to_ints(ctx.field_op(bw, "add", A, B)) == [(x + y) % p for x, y in zip(a, b)]
    assert f.limbs_to_ints(ctx.field_op(bw, "sub", A, B)) == [(x - y) % p for x, y in zip(a, b)]
    assert f.limbs_to_ints(ctx.field_op(bw, "to_mont", A)) == [x * f.r_mod % p for x in a]
    assert f.limbs_to_ints(ctx.field_op(bw, "from_mont", A)) == [x * rinv % p for x in a]
    # the prover kernels' internal arithmetic (radix 2^29 for the 255-bit field)
    assert f.limbs_to_ints(ctx.field_op(bw, "work_mul", A, B)) == [x * y * rinv % p for x, y in zip(a, b)]
    assert f.limbs_to_ints(ctx.field_op(bw, "work_fold", A, B)) == [(x + x * (y - x) * rinv) % p for x, y in zip(a, b)]


def test_bls_montgomery_product_like_reference_test(ctx):
    """tests/test_uint256.py:109-124 of the reference: encode, multiply, decode."""
    f = get_field(256)
    p = f.modulus
    rng = random.Random(381)
    a = [rng.randrange(0, p) for _ in range(32)]
    b = [rng.randrange(0, p) for _ in range(32)]
    a[:6] = [0, 1, 2, p - 1, p - 2, (1 << 255) % p]
    b[:6] = [0, 1, p - 1, 2, p - 2, 17]
    am = ctx.field_op(256, "to_mont", f.ints_to_limbs(a))
    bm = ctx.field_op(256, "to_mont", f.ints_to_limbs(b))
    out = ctx.field_op(256, "from_mont", ctx.field_op(256, "mul", am, bm))
    assert f.limbs_to_ints(out) == [x * y % p for x, y in zip(a, b)]


@pytest.mark.parametrize("bw", WIDTHS)
def test_device_transcript_matches_oracle(ctx, bw):
    f = get_field(bw)
    rng = random.Random(77 + bw)
    for npts in (2, 3, 4, 6, 8):
        for rnd in (0, 1, 13, 25):
            g = [rng.randrange(f.modulus) for _ in range(npts)]
            gm = f.ints_to_limbs([f.montgomery_encode(x) for x in g])
            r = f.limbs_to_ints(ctx.transcript_challenge(bw, rnd, gm)[None, :])[0]
            want = R.transcript_challenge(rnd, g, f.modulus, f.limbs)
            assert f.montgomery_decode(r) == want, (npts, rnd)


@pytest.mark.parametrize("bw", WIDTHS)
def test_device_inputs_match_oracle_generators(ctx, bw):
    f = get_field(bw)
    comp = Composition.from_spec(get_spec("poly_abc_plus_de"))
    with Proof(ctx, bw, comp, 9, input="affine", seed=12) as pr:
        got = pr.tables()
    want = R.affine_tables(5, 9, 12, f.modulus)
    for v in range(5):
        assert mont_ints(f, got[v]) == want[v]
    with Proof(ctx, bw, comp, 9, input="uniform", seed=99, instance=3, selector_mask=0b10010) as pr:
        got = pr.tables()
    want = R.uniform_tables(5, 9, 99, f.modulus, f.limbs, instance=3, selectors=(1, 4))
    for v in range(5):
        assert mont_ints(f, got[v]) == want[v]


# ----------------------------------------------------------------------------- proofs

@pytest.mark.parametrize("bw", WIDTHS)
def test_reference_registry_goldens(ctx, bw):
    """All 32 workloads at rounds 3 and 10, fixed and SHA3 (the reference's 384 goldens)."""
    cases = [g for g in REGISTRY if g["bit_width"] == bw]
    assert len(cases) == 96
    for g in cases:
        say("golden", g["workload"], bw, g["rounds"], g["challenge_mode"])
        comp = Composition.from_spec(get_spec(g["workload"]))
        with Proof(ctx, bw, comp, g["rounds"], input="affine", seed=1 + g["rounds"]) as pr:
            pr.run(g["challenge_mode"])
            pr.wait()
            _p, _c, cs = pr.result()
        assert 2 * cs == g["checksum_x2"], g


def _oracle_terms(comp):
    return [(c, idx) for c, idx in comp.terms()]


@pytest.mark.parametrize("bw", WIDTHS)
@pytest.mark.parametrize("num_vars", [1, 2, 5, 10])
def test_round_polys_match_oracle_uniform_sha3(ctx, bw, num_vars):
    f = get_field(bw)
    for name in ("poly_abc_plus_de", "zk_complete_add_3", "zk_jellyfish_zerocheck_hp"):
        say("polys", name, bw, num_vars)
        comp = Composition.from_spec(get_spec(name))
        tables = R.uniform_tables(comp.num_tables, num_vars, 7, f.modulus, f.limbs)
        claim, want, want_r = R.prove(tables, _oracle_terms(comp), f.modulus, num_vars,
                                      R.sha3_challenger(f.modulus, f.limbs))
        with Proof(ctx, bw, comp, num_vars, input="uniform", seed=7) as pr:
            pr.run("sha3")
            pr.wait()
            g, r, cs = pr.canonical_result()
        assert g == want
        assert r == want_r
        assert cs == R.checksum(want, f.modulus, f.limbs)
        assert (g[0][0] + g[0][1]) % f.modulus == claim


@pytest.mark.parametrize("bw", WIDTHS)
def test_random_field_coefficients_and_selectors(ctx, bw):
    """Constraint-style compositions: random field constants, selector tables in {0,1}, degree 3..6."""
    f = get_field(bw)
    rng = random.Random(500 + bw)
    T = 12
    terms = []
    for _ in range(12):
        deg = rng.randrange(3, 7)
        terms.append((rng.randrange(f.modulus), [rng.randrange(T) for _ in range(deg)]))
    comp = Composition.from_terms(T, terms)
    for nv in (3, 8):
        say("randcomp", bw, nv)
        tables = R.uniform_tables(T, nv, 31, f.modulus, f.limbs, selectors=(10, 11))
        _claim, want, want_r = R.prove(tables, _oracle_terms(comp), f.modulus, nv,
                                       R.sha3_challenger(f.modulus, f.limbs))
        with Proof(ctx, bw, comp, nv, input="uniform", seed=31, selector_mask=(1 << 10) | (1 << 11)) as pr:
            pr.run("sha3")
            pr.wait()
            g, r, _cs = pr.canonical_result()
        assert g == want and r == want_r


@pytest.mark.parametrize("bw", WIDTHS)
def test_wide_compositions_match_c_oracle(ctx, bw):
    """Interpreted compositions of more than 8 tables run as a fold launch plus a
    point-parallel evaluation launch per round: 12 tables (degree 6, 8 lanes per pair) and
    20 tables (degree 3, 4 lanes per pair), under the default schedule and with every round
    outside the single-workgroup tail, against the C oracle."""
    from oracle import cbind

    f = get_field(bw)
    rng = random.Random(900 + bw)
    for T, degs, nterms in ((12, (3, 7), 12), (20, (2, 4), 9)):
        terms = [(rng.randrange(f.modulus), [rng.randrange(T) for _ in range(rng.randrange(*degs))])
                 for _ in range(nterms)]
        terms[0] = (1, terms[0][1])
        terms[1] = (f.modulus - 1, terms[1][1])
        comp = Composition.from_terms(T, terms)
        nv = 12
        mask = (1 << (T - 1)) | (1 << (T - 2))
        want_g, want_r, _ = cbind.prove(bw, T, _oracle_terms(comp), nv, seed=41, selector_mask=mask)
        with Proof(ctx, bw, comp, nv, input="uniform", seed=41, selector_mask=mask) as pr:
            for blocks, tail in ((0, 0), (16, 1), (4096, 1)):
                say("wide", bw, T, blocks, tail)
                pr.set_schedule(blocks, tail, False)
                pr.run("sha3")
                pr.wait()
                g, r, _cs = pr.result()
                assert np.array_equal(g, want_g) and np.array_equal(r, want_r), (T, blocks, tail)


@pytest.mark.parametrize("bw", WIDTHS)
def test_point_parallel_rounds_match_c_oracle(ctx, bw):
    """The per-round kernel choice (per-pair pass kernel with the fold fused in, or a fold
    launch plus the point-parallel evaluation): no rounds, the rounds below 2^6 and 2^9
    pairs, every round, and the library default, for compiled a*b*c + d*e and for
    interpreted sums of products over up to 8 tables with field coefficients and a
    selector, against the C oracle; with fused finalize and under graph replay too."""
    from oracle import cbind

    f = get_field(bw)
    rng = random.Random(700 + bw)
    comps = [(Composition.from_spec(get_spec("poly_abc_plus_de")), 0)]
    for T, degs, nterms in ((8, (2, 6), 4), (3, (4, 8), 2), (6, (2, 4), 5)):
        terms = [(rng.randrange(f.modulus), [rng.randrange(T) for _ in range(rng.randrange(*degs))])
                 for _ in range(nterms)]
        comps.append((Composition.from_terms(T, terms), 1 << (T - 1)))
    nv = 12
    for comp, mask in comps:
        want_g, want_r, _ = cbind.prove(bw, comp.num_tables, _oracle_terms(comp), nv, seed=43,
                                        selector_mask=mask)
        with Proof(ctx, bw, comp, nv, input="uniform", seed=43, selector_mask=mask) as pr:
            for compiled in (True, False):
                pr.use_compiled(compiled)
                for wide, fused in ((0, False), (1 << 6, False), (1 << 9, True), (1 << 40, False),
                                    (-1, True)):
                    say("point-parallel", bw, comp.num_tables, compiled, wide, fused)
                    pr.set_schedule(0, 16, fused)
                    pr.set_point_parallel(wide)
                    pr.run("sha3")
                    pr.wait()
                    g, r, _cs = pr.result()
                    assert np.array_equal(g, want_g) and np.array_equal(r, want_r), (compiled, wide, fused)
            pr.set_point_parallel(1 << 9)
            pr.capture("sha3")
            for k in range(2):
                say("point-parallel graph replay", k)
                pr.run("sha3")
                pr.wait()
                g, r, _cs = pr.result()
                assert np.array_equal(g, want_g) and np.array_equal(r, want_r)
                pr.run("fixed")
                pr.wait()


@pytest.mark.parametrize("bw", WIDTHS)
def test_final_point_and_verifier(ctx, bw):
    """Final-point step: r_final follows the transcript and every table's final value is its
    multilinear extension at (r_0, ..., r_final), computed here from the input tables; the
    verifier accepts the whole proof, with claim, for compiled, interpreted and wide
    compositions, in sha3 and fixed modes, under the default schedule and with every round
    but the last outside the tail."""
    from zkduel_hip.verify import verify

    f = get_field(bw)
    p = f.modulus
    rng = random.Random(77 + bw)
    wide = Composition.from_terms(10, [(rng.randrange(p), [rng.randrange(10) for _ in range(4)]) for _ in range(6)])
    for comp, nv in ((Composition.from_spec(get_spec("poly_abc_plus_de")), 9),
                     (Composition.from_spec(get_spec("zk_complete_add_3")), 7), (wide, 8),
                     (Composition.from_spec(get_spec("poly_abc")), 1)):
        with Proof(ctx, bw, comp, nv, input="uniform", seed=13) as pr:
            tables = [[f.montgomery_decode(x) for x in f.limbs_to_ints(t)] for t in pr.tables()]
            claim = sum(_eval_terms(comp, [t[i] for t in tables], p) for i in range(1 << nv)) % p
            for mode, (blocks, tail) in (("sha3", (0, 0)), ("fixed", (0, 0)), ("sha3", (16, 1))):
                say("final", bw, comp.num_tables, nv, mode, tail)
                pr.set_schedule(blocks, tail, False)
                pr.run(mode)
                pr.wait()
                g, r, _cs = pr.result()
                rf, finals = pr.final()
                verify(bw, comp, g, r, challenge_mode=mode, claim=claim, r_final=rf, final_values=finals)
                point = [f.montgomery_decode(x) for x in f.limbs_to_ints(r)] + \
                        [f.montgomery_decode(f.limbs_to_ints(rf.reshape(1, -1))[0])]
                want = [_mle(t, point, p) for t in tables]
                assert [f.montgomery_decode(x) for x in f.limbs_to_ints(finals)] == want


def _eval_terms(comp, vals, p):
    total = 0
    for c, idx in comp.terms():
        prod = c % p
        for v in idx:
            prod = prod * vals[v] % p
        total += prod
    return total % p


def _mle(table, point, p):
    total = 0
    for i, x in enumerate(table):
        w = 1
        for k, r in enumerate(point):
            w = w * (r if (i >> k) & 1 else 1 - r) % p
        total = (total + x * w) % p
    return total


@pytest.mark.parametrize("bw", WIDTHS)
def test_package_api_matches_oracle_reference_contract(bw):
    """tests/test_zkduel.py:54-82 of the reference, at every width: explicit challenges."""
    f = get_field(bw)
    q = f.modulus
    for name in ("poly_a", "poly_ab", "poly_ab_plus_c", "poly_abc", "poly_aabbc", "poly_abc_plus_de",
                 "poly_abcg_plus_deg", "vanilla_permcheck_hp"):
        expr = get_spec(name).expression()
        rng = random.Random(256)
        rounds = 4
        names = R.expression_vars(expr)
        tables = {v: [rng.randrange(q) for _ in range(1 << rounds)] for v in names}
        rng2 = random.Random(256 + 99)
        chal = [rng2.randrange(q) for _ in range(rounds)]
        ref_claim, ref_evals = R.sumcheck_reference(tables, q=q, expression=expr, challenges=chal,
                                                    num_rounds=rounds)
        claim, evals = sumcheck(tables, q=q, expression=expr, challenges=chal, num_rounds=rounds,
                                bit_width=bw, backend="hip")
        assert claim == ref_claim
        assert evals.tolist() == ref_evals


@pytest.mark.parametrize("bw", WIDTHS)
def test_compiled_and_interpreted_compositions_agree(ctx, bw):
    """a*b, a*b*c and a*b*c + d*e have compiled kernels; forcing the interpreted ones gives
    the same bits."""
    for name, sid in (("zk_spartan_2", 1), ("poly_abc", 2), ("poly_abc_plus_de", 3)):
        comp = Composition.from_spec(get_spec(name))
        with Proof(ctx, bw, comp, 13, input="uniform", seed=17) as pr:
            assert pr.use_compiled(True) == sid
            pr.run("sha3")
            pr.wait()
            compiled = pr.result()
            pr.use_compiled(False)
            pr.run("sha3")
            pr.wait()
            interp = pr.result()
        say("compiled-vs-interpreted", name, bw)
        assert np.array_equal(compiled[0], interp[0]) and np.array_equal(compiled[1], interp[1])


@pytest.mark.parametrize("bw", [64, 256])
def test_schedule_and_graph_replays_are_identical(ctx, bw):
    comp = Composition.from_spec(get_spec("poly_abc_plus_de"))
    nv = 14
    with Proof(ctx, bw, comp, nv, input="uniform", seed=5) as pr:
        pr.run("sha3")
        pr.wait()
        base = pr.result()
        base_final = pr.final()
        for blocks, tail, fused in ((16, 1, False), (4096, 8, False), (2048, 1 << 20, False),
                                    (16, 1, True), (4096, 8, True), (0, 0, True), (0, 0, False)):
            say("schedule", bw, blocks, tail, fused)
            pr.set_schedule(blocks, tail, fused)
            pr.run("sha3")
            pr.wait()
            got = pr.result()
            assert np.array_equal(got[0], base[0]) and np.array_equal(got[1], base[1])
        # intermediate-table layouts: chunks smaller and larger than the late rounds' tables,
        # table-major, with per-pair, point-parallel and tail rounds all reading them
        pr.set_schedule(64, 8, False)
        for bad in (1, -1, 64):
            with pytest.raises(ValueError):
                pr.set_fold_chunk(bad)
        for chunk in (2, 3, 7, 11, 63, 0):
            for wide in (0, 1 << 9, -1):
                say("fold chunk", bw, chunk, wide)
                pr.set_fold_chunk(chunk)
                pr.set_point_parallel(wide)
                pr.run("sha3")
                pr.wait()
                got = pr.result()
                assert np.array_equal(got[0], base[0]) and np.array_equal(got[1], base[1])
                rf, fin = pr.final()
                assert np.array_equal(rf, base_final[0]) and np.array_equal(fin, base_final[1])
        pr.set_schedule(0, 0, False)
        pr.set_point_parallel(-1)
        pr.capture("sha3")
        for k in range(3):
            say("graph replay", k)
            pr.run("sha3")
            pr.wait()
            got = pr.result()
            assert np.array_equal(got[0], base[0]) and got[2] == base[2]
            pr.run("fixed")  # eager work between replays
            pr.wait()


@pytest.mark.parametrize("bw", [64, 256])
def test_pipelined_runs_return_each_runs_results(ctx, bw):
    """sc_proof_run_async / sc_proof_slot_result as a repeated-proving loop runs them: each run
    is enqueued before the previous run's results are read, and consecutive runs differ (SHA3
    and fixed challenges alternate), so a slot read too early or overwritten shows. Eager runs,
    then with the SHA3 runs replayed from a captured graph."""
    comp = Composition.from_spec(get_spec("poly_abc_plus_de"))
    with Proof(ctx, bw, comp, 14, input="uniform", seed=5) as pr:
        want = {}
        for mode in ("sha3", "fixed"):
            pr.run(mode)
            pr.wait()
            want[mode] = pr.result()
        assert not np.array_equal(want["sha3"][0], want["fixed"][0])
        modes = ["sha3", "fixed", "fixed", "sha3", "sha3", "fixed", "sha3", "fixed"]
        for graph in (False, True):
            if graph:
                pr.capture("sha3")
            say("pipelined", bw, "graph" if graph else "eager")
            prev = None
            for m in modes:
                slot = pr.run_async(m)
                if prev is not None:
                    got = pr.slot_result(prev[1])
                    exp = want[prev[0]]
                    assert np.array_equal(got[0], exp[0]) and np.array_equal(got[1], exp[1]) and got[2] == exp[2]
                prev = (m, slot)
            got = pr.slot_result(prev[1])
            assert np.array_equal(got[0], want[prev[0]][0]) and got[2] == want[prev[0]][2]
        with pytest.raises(ValueError):
            pr.slot_result(2)


def _sweep_cases():
    keep = []
    for g in SWEEP:
        spec = get_spec(g["workload"])
        # at-scale coverage that stays within seconds: all workloads at rounds 14, and the
        # benchmark composition at every recorded size
        if g["rounds"] == 14 or (g["workload"] == "poly_abc_plus_de" and g["rounds"] <= 26):
            keep.append(g)
    return keep


@pytest.mark.parametrize("bw", WIDTHS)
def test_reference_sweep_checksums_at_scale(ctx, bw):
    """Recorded CUDA/Triton checksums (13 runs summed) at rounds 14..26, affine, r = 1."""
    cases = [g for g in _sweep_cases() if g["bit_width"] == bw]
    assert cases
    for g in cases:
        say("sweep", g["workload"], bw, g["rounds"])
        assert g["checksum_sum"] % g["runs"] == 0
        comp = Composition.from_spec(get_spec(g["workload"]))
        with Proof(ctx, bw, comp, g["rounds"], input="affine", seed=1 + g["rounds"]) as pr:
            pr.run("fixed")
            pr.wait()
            _p, _c, cs = pr.result()
        assert cs * g["runs"] == g["checksum_sum"], g


@pytest.mark.parametrize("bw,nv", [(32, 26), (64, 26), (128, 24), (256, 24)])
def test_full_size_protocol_invariants(ctx, bw, nv):
    """The benchmark instance (a*b*c + d*e, 5 uniform tables, SHA3) at its full size."""
    f = get_field(bw)
    p = f.modulus
    comp = Composition.from_spec(get_spec("poly_abc_plus_de"))
    say("full", bw, nv)
    with Proof(ctx, bw, comp, nv, input="uniform", seed=2026) as pr:
        pr.run("sha3")
        pr.wait()
        g, r, _cs = pr.canonical_result()
    assert len(g) == nv and len(r) == nv - 1
    for j in range(nv - 1):
        assert r[j] == R.transcript_challenge(j, g[j], p, f.limbs)
    for j in range(1, nv):
        assert (g[j][0] + g[j][1]) % p == R.lagrange_eval(g[j - 1], r[j - 1], p), j


def test_autotuned_schedule_keeps_results(ctx):
    """The autotuner only changes launch schedules: the tuned proof gives the same bits."""
    from zkduel_hip import autotune

    autotune._cache.clear()
    comp = Composition.from_spec(get_spec("poly_abc_plus_de"))
    with Proof(ctx, 64, comp, 16, input="uniform", seed=8) as pr:
        pr.run("sha3")
        pr.wait()
        want = pr.result()
        choice = autotune.tune_proof(pr, repeats=1)
        say("tuned", choice["max_pass_blocks"], choice["tail_pairs"])
        pr.run("sha3")
        pr.wait()
        got = pr.result()
    assert np.array_equal(got[0], want[0]) and got[2] == want[2]
    autotune._cache.clear()
