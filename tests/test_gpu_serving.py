"""GPU tests of the continuous-batching server: every request's result is the same bits alone
and in the mix (mixed widths, sizes and compositions including requests over 8 tables, whose
steps run point-parallel, requests arriving while others run), admission respects the HBM
budget, and a Poisson replay completes with identical results."""

from __future__ import annotations

import random
import sys

import numpy as np
import pytest

from zkduel_hip import Composition, Context, Proof, get_spec
from zkduel_hip.serving import Server, run_poisson
from zkduel_hip.synthetic import random_product_sum, serving_requests

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def alone(ctx, bw, comp, nv, seed, instance, mode="sha3"):
    with Proof(ctx, bw, comp, nv, input="uniform", seed=seed, instance=instance) as pr:
        pr.run(mode)
        pr.wait()
        return pr.result()


def wide_composition(rng, tables=10):
    """More than 8 tables: a step that holds such a request runs its width's rounds
    point-parallel."""
    terms = [(1, [rng.randrange(tables) for _ in range(rng.randrange(2, 5))]) for _ in range(4)]
    terms.append((1, [tables - 1]))
    return Composition.from_terms(tables, terms)


def mixed_requests(n, seed):
    rng = random.Random(seed)
    out = []
    for i in range(n):
        bw = rng.choice((32, 64, 128, 256))
        if i % 7 == 5:
            comp = wide_composition(rng)
        elif i % 3:
            comp = random_product_sum(rng)
        else:
            comp = Composition.from_spec(get_spec("poly_abc_plus_de"))
        out.append((bw, comp, rng.randrange(1, 13), i))
    return out


def test_server_results_match_alone(ctx):
    reqs = mixed_requests(30, 1)
    with Server(ctx, 1 << 30, max_active=8) as sv:
        handles = {}
        for k, (bw, comp, nv, i) in enumerate(reqs):
            handles[sv.submit(bw, comp, nv, {"seed": 11, "instance": i})] = (bw, comp, nv, i)
            if k % 4 == 3:  # new requests arrive while earlier ones are mid-proof
                sv.step()
                sv.step()
        while any(sv.poll(h) != "done" for h in handles):
            sv.step()
        say("server steps", sv.info()["steps"])
        for h, (bw, comp, nv, i) in handles.items():
            got = sv.result(h)
            want = alone(ctx, bw, comp, nv, 11, i)
            assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1]) and got[2] == want[2], (bw, nv)


def test_two_servers_side_by_side_with_try_step(ctx):
    # two servers on one context, each on its own stream, stepped without waiting
    reqs = mixed_requests(16, 7)
    with Server(ctx, 1 << 30, max_active=8) as a, Server(ctx, 1 << 30, max_active=8) as b:
        handles = {}
        for k, (bw, comp, nv, i) in enumerate(reqs):
            sv = a if k % 2 else b
            handles[(sv, sv.submit(bw, comp, nv, {"seed": 13, "instance": i}))] = (bw, comp, nv, i)
        skipped = 0
        while any(sv.poll(h) != "done" for sv, h in handles):
            for sv in (a, b):
                skipped += not sv.try_step()[2]
        say("try_step skipped (in-flight steps full)", skipped)
        for (sv, h), (bw, comp, nv, i) in handles.items():
            got = sv.result(h)
            want = alone(ctx, bw, comp, nv, 13, i)
            assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1]) and got[2] == want[2], (bw, nv)


def test_server_budget_admits_one_at_a_time(ctx):
    comp = Composition.from_spec(get_spec("poly_abc_plus_de"))
    nv = 14  # 5 tables x 16384 x 8 B = 640 KiB input, 1.12 MiB with the fold buffers
    with Server(ctx, 3 << 19, max_active=16) as sv:
        hs = [sv.submit(64, comp, nv, {"seed": 3, "instance": i}) for i in range(5)]
        running, queued = sv.step()
        assert running == 1 and queued == 4
        while any(sv.poll(h) != "done" for h in hs):
            sv.step()
        for i, h in enumerate(hs):
            assert sv.result(h)[2] == alone(ctx, 64, comp, nv, 3, i)[2]
        with pytest.raises(MemoryError):
            sv.submit(64, comp, 20, {"seed": 3, "instance": 9})


@pytest.mark.parametrize("lanes,warm,policy", [(1, False, "all"), (2, True, "all"), (1, False, "shortest_first"),
                                               (2, False, "shortest_first")])
def test_poisson_replay_results_match_alone(ctx, lanes, warm, policy):
    # two lanes: requests over 2^14 input words run on the second server's stream, side by
    # side with the short ones; warm: the stream runs once untimed on the same servers first;
    # shortest_first: steps advance the request with the least work left first
    reqs = serving_requests(40, rate_per_s=2000.0, seed=6, nv_lo=4, nv_hi=12)
    stats = run_poisson(ctx, reqs, hbm_budget_bytes=1 << 31, max_active=16, lanes=lanes,
                        short_words=1 << 14, warm=warm, policy=policy)
    say("poisson", {k: v for k, v in stats.items() if k not in ("checksums", "latency_ms")})
    assert stats["requests"] == 40 and len(stats["checksums"]) == 40
    assert sorted(stats["latency_ms"]) == sorted(r["id"] for r in reqs)
    for r in reqs:
        want = alone(ctx, r["bit_width"], r["composition"], r["num_vars"], r["seed"], r["instance"])[2]
        assert stats["checksums"][r["id"]] == want


def test_two_lane_shortest_first_at_benchmark_sizes(ctx):
    """Requests of the serving benchmark's sizes (num_vars 17-22, widths 256/64/32, random
    product sums) on two lanes with shortest-first steps on the long one: every result equals
    the same proof run alone."""
    reqs = serving_requests(12, rate_per_s=500.0, seed=21, nv_lo=17, nv_hi=22)
    stats = run_poisson(ctx, reqs, hbm_budget_bytes=64 << 30, max_active=16, lanes=2,
                        short_words=1 << 22, policy="shortest_first")
    say("benchmark sizes", {k: v for k, v in stats.items() if k not in ("checksums", "latency_ms")},
        sorted((r["bit_width"], r["num_vars"]) for r in reqs))
    assert len(stats["checksums"]) == 12
    for r in reqs:
        want = alone(ctx, r["bit_width"], r["composition"], r["num_vars"], r["seed"], r["instance"])[2]
        assert stats["checksums"][r["id"]] == want, (r["bit_width"], r["num_vars"])


def test_request_schedules_keep_results(ctx):
    reqs = mixed_requests(16, 4)
    rng = random.Random(4)
    with Server(ctx, 1 << 30, max_active=8) as sv:
        handles = {}
        for bw, comp, nv, i in reqs:
            h = sv.submit(bw, comp, nv, {"seed": 13, "instance": i})
            sv.set_request_schedule(h, rng.choice((1, 2, 8, 32)), rng.choice((-1, 0, 1 << 40)))
            handles[h] = (bw, comp, nv, i)
        while any(sv.poll(h) != "done" for h in handles):
            sv.step()
        for h, (bw, comp, nv, i) in handles.items():
            got = sv.result(h)
            want = alone(ctx, bw, comp, nv, 13, i)
            assert np.array_equal(got[0], want[0]) and got[2] == want[2], (bw, nv)


def test_request_too_large_for_one_step_is_refused_at_submit(ctx):
    """A request whose largest round needs more partial tiles than a step holds even at the
    largest tile is refused (it could never run); one that fits with a larger tile is
    accepted. Nothing is stepped: no table memory is allocated."""
    comp = Composition.from_terms(1, [(1, [0] * 7)])  # degree 7: 32 pairs per block
    with Server(ctx, 200 * 10**9, max_active=4) as sv:
        with pytest.raises(ValueError):
            sv.submit(32, comp, 33, {"seed": 1, "instance": 0})  # 2^32 pairs in round 0
        h = sv.submit(32, comp, 31, {"seed": 1, "instance": 1})  # fits with a tile of 32
        assert sv.poll(h) == "queued"
        assert sv.info()["arena_size"] == 0  # creation and submission hold no table memory


def test_long_mixed_replay_returns_all_memory(ctx):
    """400 requests of every width at a high arrival rate under a small budget: every result
    equals the proof alone, and the arena holds nothing once all are done."""
    rng = random.Random(12)
    reqs = serving_requests(400, rate_per_s=20000.0, seed=12, nv_lo=6, nv_hi=16)
    for r in reqs:
        r["bit_width"] = rng.choice((32, 64, 128, 256))
    budget = 1 << 28
    with Server(ctx, budget, max_active=64) as sv:
        live = {}
        nxt = 0
        peak = 0
        while nxt < len(reqs) or live:
            for _ in range(8):
                if nxt < len(reqs):
                    r = reqs[nxt]
                    live[sv.submit(r["bit_width"], r["composition"], r["num_vars"],
                                   {"seed": r["seed"], "instance": r["instance"]})] = r
                    nxt += 1
            sv.step()
            info = sv.info()
            peak = max(peak, info["arena_size"])
            assert info["arena_used"] <= budget and info["arena_size"] <= budget
            for h in [h for h in live if sv.poll(h) == "done"]:
                r = live.pop(h)
                r["got"] = sv.result(h)[2]
        sv.drain()
        info = sv.info()
        say("replay", info, "peak held", peak)
        assert info["arena_used"] == 0
    for r in reqs:
        assert r["got"] == alone(ctx, r["bit_width"], r["composition"], r["num_vars"], r["seed"], r["instance"])[2], r["id"]
