"""GPU tests of batched proving: every instance of a batch must give the same bits as proving
it alone (itself checked against the oracle elsewhere), for uneven sizes and mixed
compositions, any tile size, per-pair or point-parallel rounds, and hipGraph replays; small
instances are also checked against the oracle directly."""

from __future__ import annotations

import random
import sys

import numpy as np
import pytest

from oracle import sumcheck_ref as R
from zkduel_hip import Composition, Context, Proof, get_field, get_spec
from zkduel_hip.prover import Batch
from zkduel_hip.synthetic import batch_instances, constraint_composition, random_product_sum

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def plain_result(ctx, bw, comp, nv, inp, mode):
    with Proof(ctx, bw, comp, nv, input=inp["kind"], seed=inp["seed"], instance=inp["instance"],
               selector_mask=inp.get("selector_mask", 0)) as pr:
        pr.run(mode)
        pr.wait()
        return pr.result()


@pytest.mark.parametrize("bw", [32, 64, 128, 256])
def test_batch_matches_single_proofs(ctx, bw):
    insts = batch_instances(n=24, seed=40 + bw, nv_lo=1, nv_hi=13)
    # include registry compositions (interpreted kernels) and a selector table
    insts.append((Composition.from_spec(get_spec("zk_vanilla_permcheck_hp")), 9,
                  {"kind": "uniform", "seed": 7, "instance": 100, "selector_mask": 0b10}))
    insts.append((Composition.from_spec(get_spec("poly_abc_plus_de")), 11,
                  {"kind": "affine", "seed": 12, "instance": 0}))
    if bw in (64, 256):
        # 10 tables: the batch's default rounds are point-parallel, 0 forces per-pair kernels
        wide_comp, mask = constraint_composition(bw, tables=8, selectors=2, monomials=6)
        insts.append((wide_comp, 7, {"kind": "uniform", "seed": 5, "instance": 200, "selector_mask": mask}))
    wants = {mode: [plain_result(ctx, bw, comp, nv, inp, mode) for comp, nv, inp in insts]
             for mode in ("sha3", "fixed")}
    with Batch(ctx, bw, insts) as b:
        # per-pair rounds only, point-parallel rounds at or below 2^7 pairs, all, the default
        for wide in (0, 1 << 7, 1 << 40, -1):
            b.set_point_parallel(wide)
            for items in (1, 4, 16):
                b.set_tile(items)
                for mode in ("sha3", "fixed"):
                    say("batch", bw, wide, items, mode)
                    b.run(mode)
                    b.wait()
                    for i, (comp, nv, inp) in enumerate(insts):
                        got = b.result(i)
                        want = wants[mode][i]
                        assert np.array_equal(got[0], want[0]), (i, nv, wide, items)
                        assert np.array_equal(got[1], want[1]) and got[2] == want[2]


@pytest.mark.parametrize("bw", [64, 256])
def test_batch_small_instances_match_oracle(ctx, bw):
    f = get_field(bw)
    rng = random.Random(bw)
    insts = []
    for i in range(12):
        insts.append((random_product_sum(rng), rng.randrange(1, 8), {"kind": "uniform", "seed": 3, "instance": i}))
    with Batch(ctx, bw, insts) as b:
        b.run("sha3")
        b.wait()
        for i, (comp, nv, inp) in enumerate(insts):
            tables = R.uniform_tables(comp.num_tables, nv, 3, f.modulus, f.limbs, instance=i)
            _c, want, want_r = R.prove(tables, comp.terms(), f.modulus, nv, R.sha3_challenger(f.modulus, f.limbs))
            polys, chal, _cs = b.result(i)
            got = [[f.montgomery_decode(v) for v in f.limbs_to_ints(polys[j])] for j in range(nv)]
            assert got == want
            assert [f.montgomery_decode(v) for v in f.limbs_to_ints(chal)] == want_r


def test_batch_graph_replays(ctx):
    insts = batch_instances(n=40, seed=9, nv_lo=4, nv_hi=14)
    with Batch(ctx, 64, insts) as b:
        b.set_point_parallel(1 << 8)  # both round forms in the captured stages
        b.run("sha3")
        b.wait()
        base = [b.result(i) for i in range(len(insts))]
        b.capture("sha3")
        for k in range(3):
            say("batch graph replay", k)
            b.run("sha3")
            b.wait()
            for i in range(len(insts)):
                got = b.result(i)
                assert np.array_equal(got[0], base[i][0]) and got[2] == base[i][2]


@pytest.mark.parametrize("bw", [64, 256])
def test_per_instance_schedules_keep_results(ctx, bw):
    """Random per-instance tiles, kernel choices and tail depths (sc_batch_set_schedules), and
    the per-class tuner, change no result bit; the schedule in effect reads back."""
    from zkduel_hip.autotune import tune_batch

    insts = batch_instances(n=32, seed=70 + bw, nv_lo=2, nv_hi=14)
    wants = [plain_result(ctx, bw, comp, nv, inp, "sha3") for comp, nv, inp in insts]
    rng = random.Random(bw)
    with Batch(ctx, bw, insts) as b:
        for trial in range(3):
            items = [rng.choice((1, 2, 4, 8, 16)) for _ in insts]
            wide = [rng.choice((-1, 0, 1 << 6, 1 << 40)) for _ in insts]
            tails = [rng.choice((0, 1, 16, 4096)) for _ in insts]
            b.set_schedules(items, wide, tails)
            for i in range(len(insts)):
                sched = b.instance_schedule(i)
                assert sched["items"] == items[i]
                if wide[i] >= 0:
                    assert sched["point_parallel_pairs"] == wide[i]
            say("per-instance schedules", bw, trial)
            b.run("sha3")
            b.wait()
            for i in range(len(insts)):
                got = b.result(i)
                assert np.array_equal(got[0], wants[i][0]) and np.array_equal(got[1], wants[i][1]), (trial, i)
        tuned = tune_batch(b, items=(2, 8), wide_pairs=(-1, 0), tail_pairs=(0, 16), repeats=1)
        say("tuned classes", tuned["classes"])
        assert sum(c["instances"] for c in tuned["classes"].values()) == len(insts)
        b.run("sha3")
        b.wait()
        for i in range(len(insts)):
            got = b.result(i)
            assert np.array_equal(got[0], wants[i][0]) and got[2] == wants[i][2], i
