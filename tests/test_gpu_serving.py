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


def test_poisson_replay_results_match_alone(ctx):
    reqs = serving_requests(40, rate_per_s=2000.0, seed=6, nv_lo=4, nv_hi=12)
    stats = run_poisson(ctx, reqs, hbm_budget_bytes=1 << 31, max_active=16)
    say("poisson", {k: v for k, v in stats.items() if k != "checksums"})
    assert stats["requests"] == 40 and len(stats["checksums"]) == 40
    for r in reqs:
        want = alone(ctx, r["bit_width"], r["composition"], r["num_vars"], r["seed"], r["instance"])[2]
        assert stats["checksums"][r["id"]] == want
