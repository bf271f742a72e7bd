#!/usr/bin/env python3
"""Benchmark: plain single SumCheck proofs per second on one MI355X.

Workload (one step = one whole proof, results copied to the host):
  a*b*c + d*e over 5 uniform tables from the seeded counter-based generator, SHA3 transcript
  derived on the device, Goldilocks (2^64 - 2^32 + 1) at num_vars = 26 (headline), plus the
  other three widths at their sizes (2^32-5 at 26, 2^127-1 and BLS12-381 Fr at 24) reported
  under config.per_width.

Prints ONE JSON line. `value` is the headline proof replayed as its captured hipGraph (one
host launch per proof), K timed replays after the warmups, pipelined the way a caller proving
repeatedly runs them: each proof's results are copied by the device into a pinned host slot
and read while the next proof runs (sc_proof_run_async). The same replays one at a time
(replay, wait, read) are reported as `unpipelined_proofs_per_s`, with their p50/p99. `roofline` is for the dominant kernel (the fused fold+eval pass kernel):
algorithmic bytes of every pass launch divided by their summed durations, both measured here
with hipEvents recorded on the stream the kernels run on, over K eager proofs of the same
instance (per-launch event records need eager launches); their rate is reported as
`eager_proofs_per_s`.
`cpu_baseline` times the oracle's single-core Python restatement of the reference prover
(src/zkduel/reference.py:44-110) on a bounded sample and scales it to the full size.

Usage: python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--no-cpu] [--widths 64 256]
                       [--sections widths,abc16,cfg3,batched,serving,cpu]
(--sections picks the parts to run, for profiling one of them; a run without 'widths' prints
only the parts it ran, not the benchmark line.)
"""

from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path[:0] = [str(ROOT / "sumcheck-triton_amd"), str(ROOT)]

METRIC = "SumCheck proofs/s + % HBM/INT32 roofline; batched vs plain; p50/p99 latency"
HBM_PEAK_GBPS = 8000.0
SIZES = {32: 26, 64: 26, 128: 24, 256: 24}
WORKLOAD = "poly_abc_plus_de"
SEED = 2026
SECTIONS = ("widths", "abc16", "cfg3", "batched", "serving", "cpu")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pass_bytes(T: int, limbs: int, n0: int, j: int) -> int:
    """Algorithmic HBM bytes of round j's pass: read round j-1's tables (or the input at
    j = 0) once, write round j's folded tables once."""
    w = 4 * limbs
    if j == 0:
        return T * w * n0
    return T * w * ((n0 >> (j - 1)) + (n0 >> j))


def modmuls_per_pair(spec, fold: bool) -> int:
    """Field multiplications per pair the kernel's algorithm needs in a round: the composition
    at each evaluated point (each term's len-1 products, plus one for a coefficient other than
    +-1) at t = 0..d-1 (t = 1 only in round 0: later rounds take g(1) from the claim), and the
    leading coefficient, which only the degree-d terms enter (the kernels evaluate t = d at
    infinity when d >= 2); when the round folds, two per table."""
    def products(terms):
        return sum(n - 1 + (1 if c not in (1, -1) else 0) for c, n in terms)
    terms = list(zip(spec.coeffs, spec.term_lens))
    d = spec.degree
    finite = (d - 1 if fold else d) if d >= 2 else (d if fold else d + 1)
    lead = products([(c, n) for c, n in terms if n == d]) if d >= 2 else 0
    return finite * products(terms) + lead + (2 * spec.num_vars if fold else 0)


def measure_width(ctx, bw: int, nv: int, steps: int, warmup: int, mad_rate: float | None,
                  fixed_schedule: tuple[int, int, int] | None = None) -> dict:
    from zkduel_hip import Composition, Proof, get_field, get_spec
    from zkduel_hip.driver import algorithmic_bytes

    f = get_field(bw)
    spec = get_spec(WORKLOAD)
    comp = Composition.from_spec(spec)
    T = spec.num_vars
    from zkduel_hip.autotune import tune_proof

    with Proof(ctx, bw, comp, nv, input="uniform", seed=SEED) as pr:
        if fixed_schedule:
            pr.set_schedule(fixed_schedule[0], fixed_schedule[1], False)
            pr.set_point_parallel(fixed_schedule[2])
            schedule = {"max_pass_blocks": fixed_schedule[0], "tail_pairs": fixed_schedule[1]}
            # without the autotuner's timings first, the device would start the timed proofs
            # cold: an idle MI355X ran the first ~0.1 s of proofs 10-15% slow
            t0 = time.perf_counter()
            while time.perf_counter() - t0 < 0.3:
                pr.run("sha3")
                pr.wait()
        else:
            schedule = tune_proof(pr)  # kernel choice, grid cap and tail depth timed on this shape
        # rounds at or below this pair count run fold + point-parallel evaluation, not pass_kernel
        wide_pairs = pr.info()["point_parallel_pairs"]
        schedule = {"max_pass_blocks": schedule["max_pass_blocks"], "tail_pairs": schedule["tail_pairs"],
                    "point_parallel_pairs": wide_pairs,
                    "library_default": bool(schedule.get("library_default", False)),
                    "tuned_ms": schedule.get("ms"), "default_ms": schedule.get("default_ms")}
        pr.set_profiling(True)
        for _ in range(warmup):
            pr.run("sha3")
            pr.wait()
        lat, pass_ms_sum, pass_bytes_sum, pass_launches, dev_ms = [], 0.0, 0, 0, []
        pass_macs_sum = 0
        limb_macs = 2 * f.limbs * f.limbs + f.limbs  # 32x32 multiply-accumulates per product
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            s0 = time.perf_counter()
            pr.run("sha3")
            dev_ms.append(pr.wait())
            pr.result()
            lat.append((time.perf_counter() - s0) * 1000.0)
            passes, _tail = pr.kernel_times()
            for j, ms in enumerate(passes):
                if ((1 << nv) >> (j + 1)) <= wide_pairs:
                    continue
                pass_ms_sum += ms
                pass_bytes_sum += pass_bytes(T, f.limbs, 1 << nv, j)
                pass_macs_sum += ((1 << nv) >> (j + 1)) * modmuls_per_pair(spec, j > 0) * limb_macs
                pass_launches += 1
        ctx.sync()
        wall = time.perf_counter() - t0
        info = pr.info()
        # the same proof replayed as one hipGraph (one host launch instead of one per kernel)
        pr.set_profiling(False)
        pr.capture("sha3")
        pr.run("sha3")
        pr.wait()
        ctx.sync()
        glat = []
        g0 = time.perf_counter()
        for _ in range(steps):
            s0 = time.perf_counter()
            pr.run("sha3")
            pr.wait()
            pr.result()
            glat.append((time.perf_counter() - s0) * 1000.0)
        ctx.sync()
        graph_wall = time.perf_counter() - g0
        graph_ms = graph_wall * 1000.0 / steps
        # the same replays pipelined as a caller proving repeatedly runs them: proof k + 1 is
        # enqueued before proof k's results (copied by the device into a pinned slot) are read,
        # so the host's work overlaps the device's
        prev = pr.run_async("sha3")
        pr.slot_result(prev)
        ctx.sync()
        p0 = time.perf_counter()
        prev = None
        for _ in range(steps):
            slot = pr.run_async("sha3")
            if prev is not None:
                pr.slot_result(prev)
            prev = slot
        pr.slot_result(prev)
        ctx.sync()
        pipe_wall = time.perf_counter() - p0
    lat.sort()
    glat.sort()
    achieved = pass_bytes_sum / (pass_ms_sum * 1e6) if pass_ms_sum > 0 else 0.0
    algo = algorithmic_bytes(T, f.limbs, nv)
    return {
        "bit_width": bw,
        "num_vars": nv,
        "proofs_per_s": steps / wall,
        "ms_per_proof": wall * 1000.0 / steps,
        "device_ms_median": statistics.median(dev_ms),
        "p50_ms": statistics.median(lat),
        "p99_ms": lat[min(len(lat) - 1, int(round(0.99 * (len(lat) - 1))))],
        "algo_bytes_per_proof": algo,
        "proof_hbm_gbps": algo / (statistics.median(dev_ms) * 1e6),
        "pass_kernel_gbps": achieved,
        "pass_kernel_avg_ms": pass_ms_sum / max(pass_launches, 1),
        "pass_kernel_avg_bytes": pass_bytes_sum / max(pass_launches, 1),
        "pass_rounds": info["pass_rounds"],
        "pass_kernel_rounds": sum(1 for j in range(info["pass_rounds"]) if ((1 << nv) >> (j + 1)) > wide_pairs),
        "hbm_frac": achieved / HBM_PEAK_GBPS,
        "pass_limb_macs_per_s": pass_macs_sum / (pass_ms_sum * 1e-3) if pass_ms_sum > 0 else 0.0,
        "int32_frac": (pass_macs_sum / (pass_ms_sum * 1e-3)) / mad_rate if (mad_rate and pass_ms_sum > 0) else None,
        "schedule": schedule,
        "graph_ms_per_proof": graph_ms,
        "graph_proofs_per_s": steps / graph_wall,
        "graph_p50_ms": statistics.median(glat),
        "graph_p99_ms": glat[min(len(glat) - 1, int(round(0.99 * (len(glat) - 1))))],
        "pipelined_ms_per_proof": pipe_wall * 1000.0 / steps,
        "pipelined_proofs_per_s": steps / pipe_wall,
        "wall_s": wall,
    }


def measure_plain(ctx, bw: int, comp, nv: int, reps: int, selector_mask: int = 0) -> dict:
    """Median whole-proof time of one prepared instance (SHA3, uniform tables), each proof
    replayed as one captured hipGraph."""
    from zkduel_hip import Proof

    with Proof(ctx, bw, comp, nv, input="uniform", seed=SEED, selector_mask=selector_mask) as pr:
        pr.capture("sha3")
        pr.run("sha3")
        pr.wait()
        lat = []
        for _ in range(reps):
            t0 = time.perf_counter()
            pr.run("sha3")
            pr.wait()
            pr.result()
            lat.append((time.perf_counter() - t0) * 1000.0)
    lat.sort()
    return {"bit_width": bw, "num_vars": nv, "ms_per_proof": statistics.median(lat),
            "proofs_per_s": 1000.0 / statistics.median(lat), "p99_ms": lat[-1]}


def measure_batched(ctx, bw: int, n: int = 256, reps: int = 3) -> dict:
    """Batched mix: n uneven instances (num_vars 10..22, random product sums over 2..8 tables),
    proved one at a time (plain) and as one batch; both SHA3. The batch's schedule is tuned
    per (width, degree, size) class of instance (autotune.tune_batch) before the timed runs.
    Both sides are replayed as captured hipGraphs (each plain proof its own graph), and the
    plain proofs are also timed eagerly. Results must agree."""
    from zkduel_hip import Proof
    from zkduel_hip.autotune import tune_batch
    from zkduel_hip.prover import Batch
    from zkduel_hip.synthetic import batch_instances

    insts = batch_instances(n, seed=4)
    with Batch(ctx, bw, insts) as b:
        t0 = time.perf_counter()
        tune = tune_batch(b, "sha3")
        tune_s = time.perf_counter() - t0
        b.capture("sha3")
        b.run("sha3")
        b.wait()
        t0 = time.perf_counter()
        for _ in range(reps):
            b.run("sha3")
            b.wait()
        batched_s = (time.perf_counter() - t0) / reps
        batch_cs = [b.result(i)[2] for i in range(n)]
        info = b.info()
    proofs = [Proof(ctx, bw, c, nv, input="uniform", seed=inp["seed"], instance=inp["instance"]) for c, nv, inp in insts]
    try:
        for pr in proofs:  # eager first
            pr.run("sha3")
            pr.wait()
        ctx.sync()
        t0 = time.perf_counter()
        for pr in proofs:
            pr.run("sha3")
            pr.wait()
            pr.result()
        eager_s = time.perf_counter() - t0
        for pr in proofs:
            pr.capture("sha3")
            pr.run("sha3")
            pr.wait()
        ctx.sync()
        t0 = time.perf_counter()
        plain_cs = []
        for pr in proofs:
            pr.run("sha3")
            pr.wait()
            plain_cs.append(pr.result()[2])
        plain_s = time.perf_counter() - t0
    finally:
        for pr in proofs:
            pr.close()
    return {"instances": n, "bit_width": bw,
            "tuned_classes": tune["classes"], "tune_default_ms": round(tune["default_ms"], 3),
            "tune_ms": round(tune["ms"], 3), "tune_s": round(tune_s, 2),
            "batched_proofs_per_s": n / batched_s, "plain_proofs_per_s": n / plain_s,
            "speedup": plain_s / batched_s, "batched_ms": batched_s * 1000.0, "plain_ms": plain_s * 1000.0,
            "plain": "each proof replayed as its own captured hipGraph, one after another",
            "plain_eager_ms": eager_s * 1000.0, "speedup_vs_eager_plain": eager_s / batched_s,
            "results_identical": batch_cs == plain_cs, "device_bytes": info["device_bytes"]}


def alone_latencies(ctx, reqs: list) -> tuple[dict, dict]:
    """Each request of a serving stream proved alone on the idle GPU: its captured graph
    replayed once untimed, then once timed (replay, wait, results read). Returns
    ({id: ms}, {id: checksum})."""
    from zkduel_hip import Proof

    ms: dict[int, float] = {}
    cs: dict[int, int] = {}
    for r in reqs:
        with Proof(ctx, r["bit_width"], r["composition"], r["num_vars"], input="uniform",
                   seed=r["seed"], instance=r["instance"]) as pr:
            pr.capture("sha3")
            pr.run("sha3")
            pr.wait()
            t0 = time.perf_counter()
            pr.run("sha3")
            pr.wait()
            res = pr.result()
            ms[r["id"]] = 1000.0 * (time.perf_counter() - t0)
            cs[r["id"]] = res[2]
    return ms, cs


def measure_serving(ctx, rate_per_s: float, n: int, alone: tuple[dict, dict], budget_bytes: int = 200 * 10**9,
                    schedules: dict | None = None, lanes: int = 2, warm: bool = True, policy: str = "all") -> dict:
    """Serving: n requests with Poisson arrivals at rate_per_s (num_vars log-uniform 12..24,
    widths 256/64/32 at 0.5/0.3/0.2, random product sums) replayed in real time against the
    continuous-batching server under an HBM budget (lanes = 2: a short-request and a
    long-request server side by side; policy: the long-request server's step policy; warm: the stream runs once untimed on the same servers
    first, as a long-running server has its device memory allocated). Every checksum is
    checked against the same proof run alone, and each request's latency is set beside that
    proof's own latency alone on the idle GPU (``alone``, from alone_latencies on the same
    stream): the slowdown percentiles say what sharing the GPU cost, apart from the spread of
    sizes in the mix."""
    from zkduel_hip.serving import run_poisson
    from zkduel_hip.synthetic import serving_requests

    reqs = serving_requests(n, rate_per_s=rate_per_s, seed=5)
    st = run_poisson(ctx, reqs, hbm_budget_bytes=budget_bytes, max_active=256, schedules=schedules,
                     lanes=lanes, warm=warm, policy=policy)
    alone_ms, alone_cs = alone
    lat = st["latency_ms"]
    slow = sorted(lat[i] / alone_ms[i] for i in lat)
    al = sorted(alone_ms[i] for i in lat)
    pick = lambda xs, q: xs[min(len(xs) - 1, int(round(q * (len(xs) - 1))))]
    out = {k: v for k, v in st.items() if k not in ("checksums", "latency_ms")}
    out["alone_p50_ms"] = pick(al, 0.5)
    out["alone_p99_ms"] = pick(al, 0.99)
    out["slowdown_p50"] = pick(slow, 0.5)
    out["slowdown_p99"] = pick(slow, 0.99)
    out["alone"] = "each request proved alone on the idle GPU: graph replay, wait, results read"
    out["hbm_budget_bytes"] = budget_bytes
    out["results_identical_to_alone"] = all(st["checksums"][i] == alone_cs[i] for i in lat)
    return out


def _oracle_time(bw: int, terms, T: int, nv: int, selector_mask: int = 0) -> float:
    from oracle import sumcheck_ref as R

    p, L = R.field(bw)
    sel = tuple(v for v in range(T) if (selector_mask >> v) & 1)
    tables = R.uniform_tables(T, nv, SEED, p, L, selectors=sel)
    t0 = time.perf_counter()
    R.prove(tables, terms, p, nv, R.sha3_challenger(p, L))
    return time.perf_counter() - t0


def _host_info() -> dict:
    model = ""
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "python": sys.version.split()[0]}


def _pin_one_core() -> tuple[set, int]:
    """Pins the calling thread to one CPU (the first it may run on), as `taskset -c` would;
    returns the previous affinity set and the CPU."""
    prev = os.sched_getaffinity(0)
    cpu = min(prev)
    os.sched_setaffinity(0, {cpu})
    return prev, cpu


def cpu_baseline() -> dict:
    """The oracle (single-core Python restatement of reference.py:44-110) on the benchmark
    workloads, pinned to one CPU, at the sizes of the baseline plan: a*b*c + d*e at num_vars 20
    at every width (the headline value: Goldilocks, scaled to num_vars 26 by 2^(20 - 26), the
    prover's work being linear in the table size), the Goldilocks a*b*c proof at num_vars 16 in
    full, and the BLS12-381 constraint workload at num_vars 16 (scaled to 22). Table generation
    is not timed; the SHA3 transcript is."""
    from zkduel_hip import get_spec
    from zkduel_hip.synthetic import constraint_composition

    prev, cpu = _pin_one_core()
    try:
        spec = get_spec(WORKLOAD)
        terms = [(c, list(spec.term_vars[o : o + n])) for c, n, o in zip(spec.coeffs, spec.term_lens, spec.term_offsets)]
        nv = 20
        per = {}
        for bw in (64, 32, 128, 256):
            t = _oracle_time(bw, terms, spec.num_vars, nv)
            per[f"abc_plus_de_{bw}bit"] = {"num_vars_timed": nv, "seconds": round(t, 3),
                                           "proofs_per_s_at_benchmark_size": (1.0 / t) * 2.0 ** (nv - SIZES[bw]),
                                           "benchmark_num_vars": SIZES[bw]}
        abc = get_spec("poly_abc")
        abc_terms = [(c, list(abc.term_vars[o : o + n])) for c, n, o in zip(abc.coeffs, abc.term_lens, abc.term_offsets)]
        t = _oracle_time(64, abc_terms, abc.num_vars, 16)
        per["goldilocks_abc_nv16"] = {"num_vars_timed": 16, "seconds": round(t, 3), "proofs_per_s": 1.0 / t}
        comp3, mask3 = constraint_composition(256)
        t = _oracle_time(256, list(comp3.terms()), comp3.num_tables, 16, mask3)
        per["bls_constraint_12_monomials"] = {"num_vars_timed": 16, "seconds": round(t, 3),
                                              "proofs_per_s_at_benchmark_size": (1.0 / t) * 2.0 ** (16 - 22),
                                              "benchmark_num_vars": 22}
    finally:
        os.sched_setaffinity(0, prev)
    secs = per["abc_plus_de_64bit"]["seconds"]
    full = SIZES[64]
    return {
        "value": (1.0 / secs) * 2.0 ** (nv - full),
        "unit": "proofs/s",
        "cores": 1,
        "kind": "port",
        "sample": f"oracle/sumcheck_ref.py prove(), Goldilocks a*b*c+d*e, 5 uniform tables, SHA3, "
                  f"num_vars={nv} took {secs:.2f} s on 1 core (pinned to CPU {cpu}); scaled by 2^({nv}-{full})",
        "per_workload": {k: {kk: (round(vv, 6) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                         for k, v in per.items()},
        "host": dict(_host_info(), pinned_cpu=cpu),
    }


def traffic_bytes(pass_rounds: int):
    """HBM bytes per pass-kernel launch measured by rocprofv3 FETCH_SIZE + WRITE_SIZE (committed
    under profiles/ with the command that produced it), averaged over the pass rounds this run
    used (rounds 0..pass_rounds-1), or None."""
    path = ROOT / "profiles" / "pass_kernel_traffic_goldilocks_nv26.json"
    if not path.exists():
        return None
    d = json.loads(path.read_text())
    per = d.get("traffic_bytes_per_round")
    if per and 0 < pass_rounds <= len(per):
        return int(round(sum(per[:pass_rounds]) / pass_rounds))
    return d.get("traffic_bytes_per_launch")


def _rounded(d: dict | None) -> dict | None:
    return None if d is None else {k: (round(v, 4) if isinstance(v, float) else v) for k, v in d.items()}


def measure_configs(ctx, sections: set, steps: int) -> tuple[dict | None, dict | None]:
    """The Goldilocks a*b*c num_vars 16 proof and the BLS12-381 constraint workload, each when
    its section is selected."""
    from zkduel_hip import Composition, get_spec
    from zkduel_hip.synthetic import constraint_composition

    cfg1 = cfg3 = None
    if "abc16" in sections:
        log("bench: Goldilocks a*b*c, num_vars 16")
        cfg1 = measure_plain(ctx, 64, Composition.from_spec(get_spec("poly_abc")), 16, max(5, steps))
    if "cfg3" in sections:
        log("bench: BLS12-381, 10 + 2 selector tables, 12 monomials of degree 3-6, num_vars 22")
        comp3, mask3 = constraint_composition(256)
        cfg3 = measure_plain(ctx, 256, comp3, 22, 2, selector_mask=mask3)
    return cfg1, cfg3


def measure_batched_and_serving(ctx, sections: set, batch_widths=(64, 256)) -> tuple[dict, dict]:
    """The 256-instance batched mixes (64- and 255-bit by default) and the two Poisson serving
    replays, each when its section is selected. Served requests get the batched mixes' tuned
    class schedules when the batched section ran."""
    batched = {}
    if "batched" in sections:
        for bw in batch_widths:
            log(f"bench: 256 uneven instances batched vs plain, {bw}-bit")
            batched[str(bw)] = measure_batched(ctx, bw)
            log(f"bench:   {batched[str(bw)]['batched_proofs_per_s']:.0f} vs {batched[str(bw)]['plain_proofs_per_s']:.0f} proofs/s")
    serving = {}
    if "serving" in sections:
        from zkduel_hip.synthetic import serving_requests

        schedules = {c: v for r in batched.values() for c, v in r["tuned_classes"].items()}
        n = 600
        log(f"bench: serving, the {n} requests proved alone")
        # the request sizes do not depend on the rate (one draw per arrival either way)
        alone = alone_latencies(ctx, serving_requests(n, rate_per_s=300.0, seed=5))
        # moderate load: shortest-first steps on the long lane (the smaller requests of a burst
        # finish first); the one-lane all-running server beside it; saturation: all running
        # (shortest-first gives up ~12% of saturated throughput for a lower p50)
        for key, rate, lanes, policy in (("rate_300_per_s", 300.0, 2, "shortest_first"),
                                         ("rate_300_per_s_one_lane_all_running", 300.0, 1, "all"),
                                         ("rate_5000_per_s", 5000.0, 2, "all")):
            log(f"bench: serving, Poisson arrivals at {rate:.0f}/s, {n} requests, {lanes} lane(s), {policy}")
            serving[key] = r = measure_serving(ctx, rate, n, alone, schedules=schedules or None, lanes=lanes,
                                               policy=policy)
            log(f"bench:   {r['proofs_per_s']:.0f} proofs/s, p50 {r['p50_ms']:.2f} ms, p99 {r['p99_ms']:.2f} ms, "
                f"slowdown p50 {r['slowdown_p50']:.2f} p99 {r['slowdown_p99']:.2f}")
    return batched, serving


def run_sections_only(ctx, sections: set, args) -> int:
    """A run without the per-width proofs (a profiling run of one part): prints one JSON line
    holding only the sections run; it is not the benchmark line."""
    cfg1, cfg3 = measure_configs(ctx, sections, args.steps)
    batched, serving = measure_batched_and_serving(ctx, sections, args.batch_widths)
    out = {"partial_run": sorted(sections), "goldilocks_abc_nv16": _rounded(cfg1),
           "bls_constraint_12_monomials_nv22": _rounded(cfg3),
           "batched_256_uneven_vs_plain": {bw: _rounded(r) for bw, r in batched.items()},
           "serving_poisson_mixed_widths": {k: _rounded(r) for k, r in serving.items()}}
    if "cpu" in sections:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--widths", type=int, nargs="+", default=[64, 32, 128, 256])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--pass-blocks", type=int, default=0,
                    help="fixed pass-grid cap for the per-width proofs (with --tail-pairs: skip the autotuner)")
    ap.add_argument("--tail-pairs", type=int, default=0)
    ap.add_argument("--point-parallel", type=int, default=-1,
                    help="with --pass-blocks/--tail-pairs: rounds with at most this many pairs run "
                         "point-parallel (-1: the library default)")
    ap.add_argument("--batch-widths", type=int, nargs="+", default=[64, 256], choices=[32, 64, 128, 256],
                    help="widths of the batched section's 256-instance mixes")
    ap.add_argument("--sections", default=",".join(SECTIONS),
                    help="comma-separated subset of " + ",".join(SECTIONS) + " (for profiling one part; "
                         "without 'widths' the printed line holds only the sections run, not the benchmark line)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(independent proofs shard across GPUs with no communication between them)", file=sys.stderr)
        return 2
    sections = {s.strip() for s in args.sections.split(",") if s.strip()}
    unknown = sections - set(SECTIONS)
    if unknown:
        print(f"bench.py: unknown sections {sorted(unknown)}; choose from {','.join(SECTIONS)}", file=sys.stderr)
        return 2
    if args.no_cpu:
        sections.discard("cpu")
    if args.no_batch:
        sections -= {"batched", "serving"}
    if "widths" in sections and 64 not in args.widths:
        args.widths = [64] + args.widths

    from zkduel_hip import Context

    ctx = Context(0)
    calib = ctx.calibrate()
    log("bench: calibration " + json.dumps({k: f"{v:.4g}" for k, v in calib.items()}))
    if "widths" not in sections:
        return run_sections_only(ctx, sections, args)
    per = {}
    for bw in args.widths:
        steps = args.steps if bw == 64 else max(3, args.steps // 2)
        log(f"bench: width {bw} num_vars {SIZES[bw]} steps {steps}")
        fixed = ((args.pass_blocks, args.tail_pairs, args.point_parallel)
                 if (args.pass_blocks and args.tail_pairs) else None)
        per[bw] = measure_width(ctx, bw, SIZES[bw], steps, args.warmup, calib["mad_u64_u32"], fixed)
        log(f"bench: width {bw}: {per[bw]['proofs_per_s']:.1f} proofs/s, "
            f"pass kernel {per[bw]['pass_kernel_gbps']:.0f} GB/s")
    head = per[64]
    cfg1, cfg3 = measure_configs(ctx, sections, args.steps)
    batched, serving = measure_batched_and_serving(ctx, sections, args.batch_widths)
    cpu = None
    if "cpu" in sections:
        log("bench: cpu baseline")
        cpu = cpu_baseline()
    line = {
        "metric": METRIC,
        "value": round(head["pipelined_proofs_per_s"], 3),
        "unit": "proofs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["pipelined_ms_per_proof"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic: seeded counter-based uniform tables in [0,p), SHA3 transcript challenges",
        "config": {
            "workload": "plain single SumCheck, a*b*c + d*e over 5 uniform tables, Goldilocks, "
                        "num_vars=26, degree 3, SHA3 transcript on device",
            "bit_width": 64,
            "num_vars": 26,
            "tables": 5,
            "parallelism": "single GPU",
            "step": "one whole proof replayed as its captured hipGraph, its results copied to pinned "
                    "host memory and read by the host while the next proof runs",
            "p50_ms": round(head["graph_p50_ms"], 4),
            "p99_ms": round(head["graph_p99_ms"], 4),
            "latency": "p50/p99 of one proof run alone: replay, wait, results read",
            "unpipelined_proofs_per_s": round(head["graph_proofs_per_s"], 3),
            "eager_proofs_per_s": round(head["proofs_per_s"], 3),
            "per_width": {str(bw): _rounded(r) for bw, r in per.items()},
            "goldilocks_abc_nv16": _rounded(cfg1),
            "bls_constraint_12_monomials_nv22": _rounded(cfg3),
            "batched_256_uneven_vs_plain": {bw: _rounded(r) for bw, r in batched.items()},
            "serving_poisson_mixed_widths": {k: _rounded(r) for k, r in serving.items()},
            "calibration": {
                "mad_u64_u32_lane_ops_per_s": calib["mad_u64_u32"],
                "hbm_read_gbps": calib["hbm_read"] / 1e9,
                "hbm_copy_gbps": calib["hbm_copy"] / 1e9,
                "hbm_read2_write1_gbps": calib["hbm_mix21"] / 1e9,
                "pass_kernel_frac_of_measured_copy": head["pass_kernel_gbps"] / (calib["hbm_copy"] / 1e9),
                "pass_kernel_frac_of_measured_read2_write1": head["pass_kernel_gbps"] / (calib["hbm_mix21"] / 1e9),
            },
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(head["pass_kernel_gbps"], 1),
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": round(head["pass_kernel_gbps"] / HBM_PEAK_GBPS, 4),
            "traffic": traffic_bytes(head["pass_kernel_rounds"]),
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
