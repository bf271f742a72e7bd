"""Serving: a continuous-batching proof server over the library's sc_server, and the
Poisson-arrival driver of the serving benchmark.

Requests arrive over time; the server admits them first-in-first-out while a slot and HBM
budget are free, and every step advances every running request by one round on the device
(requests of different sizes, widths and compositions share the stage launches). Latency of
a request = time its result is observed by the host - its arrival time.
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


def run_poisson(ctx: Context, requests: Sequence[dict], hbm_budget_bytes: int, max_active: int = 256,
                challenge_mode: str = "sha3", schedules: dict | None = None) -> dict:
    """Replay ``requests`` (dicts with arrival, bit_width, num_vars, composition, seed,
    instance; see synthetic.serving_requests) against a Server in real time. Returns
    throughput, latency percentiles and every request's checksum."""
    reqs = sorted(requests, key=lambda r: r["arrival"])
    checksums: dict[int, int] = {}
    latency: dict[int, float] = {}
    t_setup = time.perf_counter()
    sv = Server(ctx, hbm_budget_bytes, max_active, schedules)
    setup_s = time.perf_counter() - t_setup
    try:
        live: dict[int, dict] = {}
        nxt = 0
        t0 = time.perf_counter()
        peak_running = 0
        while nxt < len(reqs) or live:
            now = time.perf_counter() - t0
            while nxt < len(reqs) and reqs[nxt]["arrival"] <= now:
                r = reqs[nxt]
                h = sv.submit(r["bit_width"], r["composition"], r["num_vars"],
                              {"kind": "uniform", "seed": r["seed"], "instance": r["instance"]}, challenge_mode)
                live[h] = r
                nxt += 1
            if live:
                running, _queued = sv.step()
                peak_running = max(peak_running, running)
                now = time.perf_counter() - t0
                for h in [h for h in live if sv.poll(h) == "done"]:
                    r = live.pop(h)
                    latency[r["id"]] = now - r["arrival"]
                    checksums[r["id"]] = sv.result(h)[2]
            elif nxt < len(reqs):
                time.sleep(max(0.0, min(reqs[nxt]["arrival"] - now, 0.001)))
        total = time.perf_counter() - t0
        steps = sv.info()["steps"]
    finally:
        t_down = time.perf_counter()
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
        "server_setup_s": setup_s,
        "server_teardown_s": teardown_s,
        "checksums": checksums,
    }
