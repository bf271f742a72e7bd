"""Serving: a continuous-batching proof server over the library's sc_server, and the
Poisson-arrival driver of the serving benchmark.

Requests arrive over time; the server admits them first-in-first-out while a slot and HBM
budget are free, and every step advances the running requests by one round each on the device
(all of them, or with the shortest-first policy those with the least work left; requests of
different sizes, widths and compositions share the stage launches). Latency of a request =
time its result is observed by the host - its arrival time.
"""

from __future__ import annotations

import ctypes
import statistics
import time
from typing import Sequence

import numpy as np

from . import _lib
from .fields import get_field
from .prover import INPUT_KINDS, Composition, Context

STATES = {0: "queued", 1: "running", 2: "done"}
POLICIES = {"all": 0, "shortest_first": 1}


class Server:
    """``schedules``: optional per-class schedules, {class key: {"items", "max_pairs"}} with keys
    from autotune.instance_class (as autotune.tune_batch returns them under "classes"); a
    submitted request of a listed class gets that schedule."""

    def __init__(self, ctx: Context, hbm_budget_bytes: int, max_active: int = 256,
                 schedules: dict | None = None):
        self.ctx = ctx
        self.lib = ctx.lib
        self.schedules = dict(schedules or {})
        handle = ctypes.c_void_p()
        _lib.check(self.lib.sc_server_create(ctx.handle, int(hbm_budget_bytes), int(max_active), ctypes.byref(handle)))
        self.handle = handle
        self._shape: dict[int, tuple[int, int, int]] = {}

    def close(self) -> None:
        if self.handle:
            self.lib.sc_server_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __enter__(self) -> "Server":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def submit(self, bit_width: int, comp: Composition, num_vars: int, input: dict | None = None,
               challenge_mode: str = "sha3") -> int:
        inp = input or {}
        f = get_field(bit_width)
        cs, keep = comp.bind(f)
        si = _lib.ScInput(INPUT_KINDS[inp.get("kind", "uniform")], int(inp.get("instance", 0)) & 0xFFFFFFFF, None,
                          int(inp.get("seed", 1)) & ((1 << 64) - 1), int(inp.get("selector_mask", 0)) & ((1 << 64) - 1))
        h = ctypes.c_int64()
        _lib.check(self.lib.sc_server_submit(self.handle, bit_width, ctypes.byref(cs), int(num_vars), ctypes.byref(si),
                                             _lib.CHALLENGE_MODES[challenge_mode], ctypes.byref(h)))
        del keep
        self._shape[h.value] = (bit_width, int(num_vars), comp.degree + 1)
        if self.schedules:
            from .autotune import instance_class

            sched = self.schedules.get(instance_class(bit_width, comp, int(num_vars)))
            if sched is not None:
                self.set_request_schedule(h.value, sched.get("items", 0), sched.get("max_pairs", -1))
        return h.value

    def set_request_schedule(self, handle: int, items: int = 0, max_pairs: int = -1) -> None:
        """Tile (pairs per thread; 0 keeps) and point-parallel threshold (negative: default) of a
        queued or running request. Results do not depend on them."""
        _lib.check(self.lib.sc_server_set_request_schedule(self.handle, int(handle), int(items), int(max_pairs)))

    def step(self) -> tuple[int, int]:
        running, queued = ctypes.c_int32(), ctypes.c_int32()
        _lib.check(self.lib.sc_server_step(self.handle, ctypes.byref(running), ctypes.byref(queued)))
        return running.value, queued.value

    def set_policy(self, policy: str) -> None:
        """"all": each step advances every running request by one round (the default);
        "shortest_first": the request with the least work left, every request whose current
        round costs at most a quarter of that one's, and further requests in order of work
        left while the step holds about 0.25 ms of device work. Results do not depend on it."""
        _lib.check(self.lib.sc_server_set_policy(self.handle, POLICIES[policy]))

    def try_step(self) -> tuple[int, int, bool]:
        """step() that does not wait for in-flight steps: with the server's in-flight steps
        already enqueued it enqueues nothing (sc_server_try_step). Returns (running, queued,
        stepped)."""
        running, queued, stepped = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        _lib.check(self.lib.sc_server_try_step(self.handle, ctypes.byref(running), ctypes.byref(queued),
                                               ctypes.byref(stepped)))
        return running.value, queued.value, bool(stepped.value)

    def poll(self, handle: int) -> str:
        st = ctypes.c_int32()
        _lib.check(self.lib.sc_server_poll(self.handle, int(handle), ctypes.byref(st)))
        return STATES[st.value]

    def drain(self) -> None:
        _lib.check(self.lib.sc_server_drain(self.handle))

    def result(self, handle: int) -> tuple[np.ndarray, np.ndarray, int]:
        bw, nv, npts = self._shape.pop(handle)
        L = get_field(bw).limbs
        polys = np.zeros((nv, npts, L), dtype=np.uint32)
        chal = np.zeros((max(nv - 1, 1), L), dtype=np.uint32)
        cs = ctypes.c_uint64()
        _lib.check(self.lib.sc_server_result(self.handle, int(handle), _lib.as_u32p(polys), _lib.as_u32p(chal),
                                             ctypes.byref(cs)))
        return polys, chal[: nv - 1], int(cs.value)

    def info(self) -> dict:
        used, size, steps = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        _lib.check(self.lib.sc_server_info(self.handle, ctypes.byref(used), ctypes.byref(size), ctypes.byref(steps)))
        return {"arena_used": used.value, "arena_size": size.value, "steps": steps.value}


SHORT_WORDS = 1 << 24  # input words (tables x 2^num_vars x limbs) of a short request: 64 MiB


def request_words(bit_width: int, comp: Composition, num_vars: int) -> int:
    """Input size of a request in 32-bit words: what its first round reads."""
    return comp.num_tables * (1 << int(num_vars)) * get_field(bit_width).limbs


def run_poisson(ctx: Context, requests: Sequence[dict], hbm_budget_bytes: int, max_active: int = 256,
                challenge_mode: str = "sha3", schedules: dict | None = None, lanes: int = 1,
                short_words: int = SHORT_WORDS, warm: bool = False, policy: str = "all") -> dict:
    """Replay ``requests`` (dicts with arrival, bit_width, num_vars, composition, seed,
    instance; see synthetic.serving_requests) against a server in real time. Returns
    throughput, latency percentiles, every request's latency (ms, by request id) and every
    request's checksum.

    lanes = 2 runs two servers side by side on the device, each on its own stream: requests
    of at most ``short_words`` input words go to the short lane (an eighth of the HBM budget),
    the rest to the long lane, so a short request's rounds do not queue behind a long
    request's. The host loop steps each lane that has room (Server.try_step), without waiting
    for either lane's in-flight steps. ``policy`` is the long-request server's step policy
    (Server.set_policy; with one lane, the server's); the short lane keeps "all". warm = True
    first runs the whole request stream once, all submitted at once and untimed, on the same
    servers: the timed replay then sees a server whose device memory is already allocated, as
    a long-running one is."""
    if lanes not in (1, 2):
        raise ValueError("lanes must be 1 or 2")
    if policy not in POLICIES:
        raise ValueError(f"policy must be one of {sorted(POLICIES)}")
    reqs = sorted(requests, key=lambda r: r["arrival"])
    checksums: dict[int, int] = {}
    latency: dict[int, float] = {}
    t_setup = time.perf_counter()
    if lanes == 1:
        servers = [Server(ctx, hbm_budget_bytes, max_active, schedules)]
    else:
        short_budget = hbm_budget_bytes // 8
        servers = [Server(ctx, short_budget, max_active, schedules),
                   Server(ctx, hbm_budget_bytes - short_budget, max_active, schedules)]
    servers[-1].set_policy(policy)
    setup_s = time.perf_counter() - t_setup

    def lane_of(r: dict) -> int:
        if lanes == 1:
            return 0
        return 0 if request_words(r["bit_width"], r["composition"], r["num_vars"]) <= short_words else 1

    def submit(r: dict) -> tuple[int, int]:
        k = lane_of(r)
        return k, servers[k].submit(r["bit_width"], r["composition"], r["num_vars"],
                                    {"kind": "uniform", "seed": r["seed"], "instance": r["instance"]},
                                    challenge_mode)

    def step_all(live_lanes: set) -> int:
        if lanes == 1:
            return servers[0].step()[0]
        running, stepped = 0, False
        for k in live_lanes:
            n, _q, did = servers[k].try_step()
            running += n
            stepped = stepped or did
        if not stepped:
            time.sleep(2e-5)
        return running

    try:
        warm_s = 0.0
        if warm and reqs:
            w0 = time.perf_counter()
            pending = [submit(r) for r in reqs]
            while pending:
                step_all({k for k, _ in pending})
                still = []
                for k, h in pending:
                    if servers[k].poll(h) == "done":
                        servers[k].result(h)
                    else:
                        still.append((k, h))
                pending = still
            warm_s = time.perf_counter() - w0
        steps0 = sum(sv.info()["steps"] for sv in servers)
        live: dict[tuple[int, int], dict] = {}
        nxt = 0
        t0 = time.perf_counter()
        peak_running = 0
        while nxt < len(reqs) or live:
            now = time.perf_counter() - t0
            while nxt < len(reqs) and reqs[nxt]["arrival"] <= now:
                r = reqs[nxt]
                live[submit(r)] = r
                nxt += 1
            if live:
                running = step_all({k for k, _ in live})
                peak_running = max(peak_running, running)
                now = time.perf_counter() - t0
                for key in [key for key in live if servers[key[0]].poll(key[1]) == "done"]:
                    r = live.pop(key)
                    latency[r["id"]] = now - r["arrival"]
                    checksums[r["id"]] = servers[key[0]].result(key[1])[2]
            elif nxt < len(reqs):
                time.sleep(max(0.0, min(reqs[nxt]["arrival"] - now, 0.001)))
        total = time.perf_counter() - t0
        steps = sum(sv.info()["steps"] for sv in servers) - steps0
    finally:
        t_down = time.perf_counter()
        for sv in servers:
            sv.close()
        teardown_s = time.perf_counter() - t_down
    lat = sorted(latency.values())
    return {
        "requests": len(reqs),
        "wall_s": total,
        "proofs_per_s": len(reqs) / total,
        "offered_rate_per_s": len(reqs) / reqs[-1]["arrival"] if reqs else 0.0,
        "p50_ms": 1000.0 * statistics.median(lat),
        "p99_ms": 1000.0 * lat[min(len(lat) - 1, int(round(0.99 * (len(lat) - 1))))],
        "peak_running": peak_running,
        "steps": steps,
        "lanes": lanes,
        "policy": policy,
        "warm": warm,
        "warm_s": warm_s,
        "server_setup_s": setup_s,
        "server_teardown_s": teardown_s,
        "latency_ms": {i: 1000.0 * v for i, v in latency.items()},
        "checksums": checksums,
    }
