"""Python host objects over the C ABI: compositions, a device context and prepared proofs.

A ``Proof`` is the reference's per-case state (cuda/src/field_sumcheck.cu:1178-1221): tables
prepared once on the device, then run any number of times (first run, warmups, repeats),
each run enqueuing the whole round loop with the transcript computed on the GPU.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Sequence

import numpy as np

from . import _lib
from .expr import Expression, encode_expression_with_coeffs, expression_vars
from .fields import FieldSpec, get_field
from .workloads import WorkloadSpec

INPUT_KINDS = {
    "host-canonical": _lib.SC_INPUT_HOST_CANONICAL,
    "host-mont": _lib.SC_INPUT_HOST_MONT,
    "device-mont": _lib.SC_INPUT_DEVICE_MONT,
    "affine": _lib.SC_INPUT_AFFINE,
    "uniform": _lib.SC_INPUT_UNIFORM,
    "host-mont-limb": _lib.SC_INPUT_HOST_MONT_LIMB,
    "device-mont-limb": _lib.SC_INPUT_DEVICE_MONT_LIMB,
}


@dataclass(frozen=True)
class Composition:
    """sum_k coeffs[k] * prod of tables term_vars[term_offsets[k] : +term_lens[k]].

    Coefficients are integers (negative allowed) or field elements; they are reduced mod p
    when the composition is bound to a field.
    """

    num_tables: int
    coeffs: tuple[int, ...]
    term_lens: tuple[int, ...]
    term_offsets: tuple[int, ...]
    term_vars: tuple[int, ...]

    @property
    def degree(self) -> int:
        return max(self.term_lens)

    @property
    def num_terms(self) -> int:
        return len(self.term_lens)

    @classmethod
    def from_spec(cls, spec: WorkloadSpec) -> "Composition":
        return cls(spec.num_vars, spec.coeffs, spec.term_lens, spec.term_offsets, spec.term_vars)

    @classmethod
    def from_expression(cls, expression: Expression) -> tuple["Composition", list[str]]:
        names, coeffs, lens, offs, flat = encode_expression_with_coeffs(expression)
        return cls(len(names), coeffs, lens, offs, flat), names

    @classmethod
    def from_terms(cls, num_tables: int, terms: Sequence[tuple[int, Sequence[int]]]) -> "Composition":
        coeffs, lens, offs, flat = [], [], [], []
        for c, idx in terms:
            offs.append(len(flat))
            coeffs.append(int(c))
            lens.append(len(idx))
            flat.extend(int(i) for i in idx)
        return cls(num_tables, tuple(coeffs), tuple(lens), tuple(offs), tuple(flat))

    def terms(self) -> list[tuple[int, list[int]]]:
        return [
            (c, list(self.term_vars[o : o + n]))
            for c, n, o in zip(self.coeffs, self.term_lens, self.term_offsets)
        ]

    def bind(self, field: FieldSpec):
        """The C struct plus the arrays it points into (keep both alive)."""
        lens = np.asarray(self.term_lens, dtype=np.int32)
        offs = np.asarray(self.term_offsets, dtype=np.int32)
        flat = np.asarray(self.term_vars, dtype=np.int32)
        coeffs = field.ints_to_limbs([c % field.modulus for c in self.coeffs])
        s = _lib.ScComposition(
            self.num_tables,
            self.num_terms,
            lens.ctypes.data_as(_lib.i32p),
            offs.ctypes.data_as(_lib.i32p),
            flat.ctypes.data_as(_lib.i32p),
            coeffs.ctypes.data_as(_lib.u32p),
        )
        return s, (lens, offs, flat, coeffs)


class Context:
    """One device and one HIP stream (sc_context)."""

    def __init__(self, device: int = 0):
        self.lib = _lib.load()
        handle = ctypes.c_void_p()
        _lib.check(self.lib.sc_context_create(int(device), ctypes.byref(handle)))
        self.handle = handle
        self.device = int(device)

    def close(self) -> None:
        if self.handle:
            self.lib.sc_context_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __enter__(self) -> "Context":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def sync(self) -> None:
        _lib.check(self.lib.sc_context_sync(self.handle))

    def field_op(self, bit_width: int, op: str, a: np.ndarray, b: np.ndarray | None = None) -> np.ndarray:
        """Elementwise device op on [n][L] limb arrays: mul (Montgomery), add, sub, to_mont, from_mont."""
        ops = {"mul": 0, "add": 1, "sub": 2, "to_mont": 3, "from_mont": 4, "work_mul": 5, "work_fold": 6}
        f = get_field(bit_width)
        a = np.ascontiguousarray(a, dtype=np.uint32).reshape(-1, f.limbs)
        bb = None if b is None else np.ascontiguousarray(b, dtype=np.uint32).reshape(-1, f.limbs)
        out = np.empty_like(a)
        _lib.check(self.lib.sc_field_op(self.handle, f.bit_width, ops[op], a.shape[0],
                                        _lib.as_u32p(a), _lib.as_u32p(bb), _lib.as_u32p(out)))
        return out

    def calibrate(self) -> dict[str, float]:
        """Measured lane rates (ops/s) of mad_u64_u32, mul_hi_u32, 64-bit add and fma_f64, and
        streaming HBM bandwidth (bytes/s): read-only, copy, and 2 bytes read per byte written."""
        rates = (ctypes.c_double * 7)()
        _lib.check(self.lib.sc_calibrate(self.handle, rates, 7))
        keys = ("mad_u64_u32", "mul_hi_u32", "add64", "fma_f64", "hbm_read", "hbm_copy", "hbm_mix21")
        return dict(zip(keys, list(rates)))

    def transcript_challenge(self, bit_width: int, round_idx: int, g_mont: np.ndarray) -> np.ndarray:
        """Device-derived challenge (Montgomery limbs) for g given as [npts][L] Montgomery limbs."""
        f = get_field(bit_width)
        g = np.ascontiguousarray(g_mont, dtype=np.uint32).reshape(-1, f.limbs)
        out = np.empty(f.limbs, dtype=np.uint32)
        _lib.check(self.lib.sc_transcript_challenge(self.handle, f.bit_width, int(round_idx),
                                                    _lib.as_u32p(g), g.shape[0], _lib.as_u32p(out)))
        return out


_default_ctx: Context | None = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


class Proof:
    """A prepared SumCheck instance on the device (sc_proof)."""

    def __init__(
        self,
        ctx: Context,
        bit_width: int,
        comp: Composition,
        num_vars: int,
        *,
        input: str = "uniform",
        data: np.ndarray | int | None = None,
        seed: int = 1,
        instance: int = 0,
        selector_mask: int = 0,
    ):
        self.ctx = ctx
        self.lib = ctx.lib
        self.field = get_field(bit_width)
        self.comp = comp
        self.num_vars = int(num_vars)
        self.npts = comp.degree + 1
        if input not in INPUT_KINDS:
            raise ValueError(f"input must be one of {sorted(INPUT_KINDS)}")
        kind = INPUT_KINDS[input]
        self._keep = None
        ptr = None
        if kind in (_lib.SC_INPUT_HOST_CANONICAL, _lib.SC_INPUT_HOST_MONT, _lib.SC_INPUT_HOST_MONT_LIMB):
            arr = np.ascontiguousarray(data, dtype=np.uint32)
            expect = comp.num_tables * (1 << self.num_vars) * self.field.limbs
            if arr.size != expect:
                raise ValueError(f"input has {arr.size} words, expected {expect}")
            self._keep = arr
            ptr = arr.ctypes.data
        elif kind in (_lib.SC_INPUT_DEVICE_MONT, _lib.SC_INPUT_DEVICE_MONT_LIMB):
            ptr = int(data)
        inp = _lib.ScInput(kind, int(instance) & 0xFFFFFFFF, ptr, int(seed) & ((1 << 64) - 1),
                           int(selector_mask) & ((1 << 64) - 1))
        cs, self._comp_keep = comp.bind(self.field)
        handle = ctypes.c_void_p()
        _lib.check(self.lib.sc_proof_create(ctx.handle, self.field.bit_width, ctypes.byref(cs),
                                            self.num_vars, ctypes.byref(inp), ctypes.byref(handle)))
        self.handle = handle
        self._keep = None  # host input has been copied to the device
        self.compiled = True  # matching compiled kernels in use (use_compiled)

    def close(self) -> None:
        if self.handle:
            self.lib.sc_proof_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __enter__(self) -> "Proof":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _explicit(self, mode: int, challenges) -> np.ndarray | None:
        if mode != _lib.SC_CHALLENGE_EXPLICIT:
            return None
        vals = [int(c) % self.field.modulus for c in list(challenges)[: max(self.num_vars - 1, 0)]]
        if len(vals) < self.num_vars - 1:
            raise ValueError(f"need {self.num_vars - 1} challenges, got {len(vals)}")
        if not vals:
            return np.zeros((1, self.field.limbs), dtype=np.uint32)
        return self.field.ints_to_limbs(vals)

    def set_schedule(self, max_pass_blocks: int = 0, tail_pairs: int = 0, fused_finalize: bool = False) -> None:
        _lib.check(self.lib.sc_proof_set_schedule(self.handle, int(max_pass_blocks), int(tail_pairs),
                                                  1 if fused_finalize else 0))

    def set_point_parallel(self, max_pairs: int = -1) -> None:
        """Rounds with at most ``max_pairs`` pairs run as fold + point-parallel evaluation
        instead of the per-pair pass kernel (0: never; negative: the default rule)."""
        _lib.check(self.lib.sc_proof_set_point_parallel(self.handle, int(max_pairs)))

    def set_fold_chunk(self, chunk_log2: int = 0) -> None:
        """Intermediate tables in chunks of 2^chunk_log2 entries (chunk_log2 >= 2) of every table side by side
        (0: the default, 2^7 up to 128 bits and table-major at 256 bits; a chunk at least a
        table's size: table-major). Results do not depend on it."""
        _lib.check(self.lib.sc_proof_set_fold_chunk(self.handle, int(chunk_log2)))

    def run(self, challenge_mode: str = "sha3", challenges=None) -> None:
        mode = _lib.CHALLENGE_MODES[challenge_mode]
        r = self._explicit(mode, challenges)
        _lib.check(self.lib.sc_proof_run(self.handle, mode, _lib.as_u32p(r)))

    def capture(self, challenge_mode: str = "sha3", challenges=None) -> None:
        mode = _lib.CHALLENGE_MODES[challenge_mode]
        r = self._explicit(mode, challenges)
        _lib.check(self.lib.sc_proof_capture(self.handle, mode, _lib.as_u32p(r)))

    def wait(self) -> float:
        ms = ctypes.c_float()
        _lib.check(self.lib.sc_proof_wait(self.handle, ctypes.byref(ms)))
        return float(ms.value)

    def result(self) -> tuple[np.ndarray, np.ndarray, int]:
        """(round_polys [nv][d+1][L], challenges [nv-1][L], checksum), Montgomery limbs."""
        L = self.field.limbs
        polys = np.zeros((self.num_vars, self.npts, L), dtype=np.uint32)
        chal = np.zeros((max(self.num_vars - 1, 1), L), dtype=np.uint32)
        cs = ctypes.c_uint64()
        _lib.check(self.lib.sc_proof_result(self.handle, _lib.as_u32p(polys), _lib.as_u32p(chal),
                                            ctypes.byref(cs)))
        return polys, chal[: self.num_vars - 1], int(cs.value)

    def run_async(self, challenge_mode: str = "sha3", challenges=None) -> int:
        """Enqueues the proof and a copy of its results into a pinned host slot; returns the
        slot for slot_result. A slot is reused two runs later (read it before then)."""
        mode = _lib.CHALLENGE_MODES[challenge_mode]
        r = self._explicit(mode, challenges)
        slot = ctypes.c_int32()
        _lib.check(self.lib.sc_proof_run_async(self.handle, mode, _lib.as_u32p(r), ctypes.byref(slot)))
        return int(slot.value)

    def slot_result(self, slot: int) -> tuple[np.ndarray, np.ndarray, int]:
        """result() of the run that filled slot, waiting for that run's copy only."""
        L = self.field.limbs
        polys = np.zeros((self.num_vars, self.npts, L), dtype=np.uint32)
        chal = np.zeros((max(self.num_vars - 1, 1), L), dtype=np.uint32)
        cs = ctypes.c_uint64()
        _lib.check(self.lib.sc_proof_slot_result(self.handle, int(slot), _lib.as_u32p(polys), _lib.as_u32p(chal),
                                                 ctypes.byref(cs)))
        return polys, chal[: self.num_vars - 1], int(cs.value)

    def tables(self) -> np.ndarray:
        out = np.empty((self.comp.num_tables, 1 << self.num_vars, self.field.limbs), dtype=np.uint32)
        _lib.check(self.lib.sc_proof_tables(self.handle, _lib.as_u32p(out)))
        return out

    def info(self) -> dict:
        d, pr, nb, wp = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_uint64(), ctypes.c_int64()
        _lib.check(self.lib.sc_proof_info(self.handle, ctypes.byref(d), ctypes.byref(pr), ctypes.byref(nb),
                                          ctypes.byref(wp)))
        return {"degree": d.value, "pass_rounds": pr.value, "device_bytes": nb.value,
                "point_parallel_pairs": wp.value}

    def use_compiled(self, enable: bool = True) -> int:
        """Use the compiled kernels of a matching composition (default) or force the
        interpreted ones. Returns the matching compiled composition id (0: none)."""
        rc = self.lib.sc_proof_use_compiled(self.handle, 1 if enable else 0)
        if rc < 0:
            _lib.check(-rc)
        self.compiled = bool(enable)
        return rc

    def set_profiling(self, enable: bool = True) -> None:
        _lib.check(self.lib.sc_proof_set_profiling(self.handle, 1 if enable else 0))

    def final(self) -> tuple[np.ndarray, np.ndarray]:
        """Final-point step of the last run: (r_final [L], final_values [T][L]), Montgomery
        limbs: the challenge derived from the last round polynomial and every table folded
        at the full challenge point. Needs a run with fixed or sha3 challenges."""
        L = self.field.limbs
        r = np.zeros(L, dtype=np.uint32)
        vals = np.zeros((self.comp.num_tables, L), dtype=np.uint32)
        _lib.check(self.lib.sc_proof_final(self.handle, _lib.as_u32p(r), _lib.as_u32p(vals)))
        return r, vals

    def kernel_times(self) -> tuple[list[float], float]:
        """(pass kernel ms per fused round, tail kernel ms) of the last profiled run."""
        rounds = self.info()["pass_rounds"]
        buf = (ctypes.c_float * max(rounds, 1))()
        tail = ctypes.c_float()
        _lib.check(self.lib.sc_proof_kernel_times(self.handle, buf, rounds, ctypes.byref(tail)))
        return [float(buf[j]) for j in range(rounds)], float(tail.value)

    def round_times(self) -> tuple[list[float], list[float], float]:
        """(pass ms per round before the tail, finalize ms per such round, tail ms) of the last
        profiled run."""
        rounds = self.info()["pass_rounds"]
        passes = (ctypes.c_float * max(rounds, 1))()
        fins = (ctypes.c_float * max(rounds, 1))()
        tail = ctypes.c_float()
        _lib.check(self.lib.sc_proof_round_times(self.handle, passes, fins, rounds, ctypes.byref(tail)))
        return [float(passes[j]) for j in range(rounds)], [float(fins[j]) for j in range(rounds)], float(tail.value)

    def canonical_result(self) -> tuple[list[list[int]], list[int], int]:
        """Round polynomials and challenges decoded to canonical ints, plus the checksum."""
        polys, chal, cs = self.result()
        f = self.field
        g = [[f.montgomery_decode(v) for v in f.limbs_to_ints(polys[j])] for j in range(self.num_vars)]
        r = [f.montgomery_decode(v) for v in f.limbs_to_ints(chal)] if self.num_vars > 1 else []
        return g, r, cs


class Batch:
    """Many independent instances proved together, round by round (sc_batch).

    ``instances`` is a list of ``(composition, num_vars, input)`` where input is a dict for
    the generator: ``{"kind": "uniform"|"affine", "seed": int, "instance": int,
    "selector_mask": int}``.
    """

    def __init__(self, ctx: Context, bit_width: int, instances: Sequence[tuple]):
        self.ctx = ctx
        self.lib = ctx.lib
        self.field = get_field(bit_width)
        n = len(instances)
        if n == 0:
            raise ValueError("a batch needs at least one instance")
        self.comps = [c for c, _nv, _inp in instances]
        self.num_vars = [int(nv) for _c, nv, _inp in instances]
        bound = [c.bind(self.field) for c in self.comps]
        self._keep = [k for _s, k in bound]
        carr = (_lib.ScComposition * n)(*[s for s, _k in bound])
        nvs = np.asarray(self.num_vars, dtype=np.int32)
        inps = (_lib.ScInput * n)()
        for i, (_c, _nv, inp) in enumerate(instances):
            kind = INPUT_KINDS[inp.get("kind", "uniform")]
            inps[i] = _lib.ScInput(kind, int(inp.get("instance", i)) & 0xFFFFFFFF, None,
                                   int(inp.get("seed", 1)) & ((1 << 64) - 1),
                                   int(inp.get("selector_mask", 0)) & ((1 << 64) - 1))
        handle = ctypes.c_void_p()
        _lib.check(self.lib.sc_batch_create(ctx.handle, self.field.bit_width, n, carr,
                                            nvs.ctypes.data_as(_lib.i32p), inps, ctypes.byref(handle)))
        self.handle = handle

    def autotune(self, candidates: Sequence[int] = (1, 2, 4, 8, 16, 32), challenge_mode: str = "sha3",
                 repeats: int = 1) -> dict[int, float]:
        """Times one batch run per tile size (pairs per thread) after a warm-up run and keeps
        the fastest; returns {items: ms}. The tile changes only the schedule, never the
        results."""
        times: dict[int, float] = {}
        for items in candidates:
            self.set_tile(items)
            self.run(challenge_mode)
            self.wait()
            best = float("inf")
            for _ in range(max(1, repeats)):
                self.run(challenge_mode)
                best = min(best, self.wait())
            times[int(items)] = best
        self.set_tile(min(times, key=times.get))
        return times

    def close(self) -> None:
        if self.handle:
            self.lib.sc_batch_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __enter__(self) -> "Batch":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self) -> int:
        return len(self.comps)

    def set_tile(self, items_per_thread: int) -> None:
        _lib.check(self.lib.sc_batch_set_tile(self.handle, int(items_per_thread)))

    def set_point_parallel(self, max_pairs: int = -1) -> None:
        """Instance rounds with at most ``max_pairs`` pairs run point-parallel (0: never;
        negative: the default rule)."""
        _lib.check(self.lib.sc_batch_set_point_parallel(self.handle, int(max_pairs)))

    def set_schedules(self, items: Sequence[int] | None = None, max_pairs: Sequence[int] | None = None,
                      tail_pairs: Sequence[int] | None = None) -> None:
        """Per-instance schedules, one value per instance (None keeps the current ones): tile
        (pairs per thread; 0 keeps), point-parallel threshold (negative: default) and tail
        threshold in pairs (0: default). Results do not depend on them."""
        n = len(self.comps)

        def arr(vals, dtype):
            if vals is None:
                return None
            a = np.ascontiguousarray(vals, dtype=dtype)
            if a.shape != (n,):
                raise ValueError(f"need {n} values, one per instance")
            return a

        it, mp, tp = arr(items, np.int32), arr(max_pairs, np.int64), arr(tail_pairs, np.int64)
        i64p = ctypes.POINTER(ctypes.c_int64)
        _lib.check(self.lib.sc_batch_set_schedules(
            self.handle, it.ctypes.data_as(_lib.i32p) if it is not None else None,
            mp.ctypes.data_as(i64p) if mp is not None else None, tp.ctypes.data_as(i64p) if tp is not None else None))

    def instance_schedule(self, i: int) -> dict:
        """{"items", "point_parallel_pairs", "tail_round"} of instance i, as in effect."""
        it, mp, tr = ctypes.c_int32(), ctypes.c_int64(), ctypes.c_int32()
        _lib.check(self.lib.sc_batch_instance_schedule(self.handle, int(i), ctypes.byref(it), ctypes.byref(mp),
                                                       ctypes.byref(tr)))
        return {"items": it.value, "point_parallel_pairs": mp.value, "tail_round": tr.value}

    def run(self, challenge_mode: str = "sha3") -> None:
        _lib.check(self.lib.sc_batch_run(self.handle, _lib.CHALLENGE_MODES[challenge_mode]))

    def capture(self, challenge_mode: str = "sha3") -> None:
        _lib.check(self.lib.sc_batch_capture(self.handle, _lib.CHALLENGE_MODES[challenge_mode]))

    def wait(self) -> float:
        ms = ctypes.c_float()
        _lib.check(self.lib.sc_batch_wait(self.handle, ctypes.byref(ms)))
        return float(ms.value)

    def result(self, i: int) -> tuple[np.ndarray, np.ndarray, int]:
        L = self.field.limbs
        nv = self.num_vars[i]
        polys = np.zeros((nv, self.comps[i].degree + 1, L), dtype=np.uint32)
        chal = np.zeros((max(nv - 1, 1), L), dtype=np.uint32)
        cs = ctypes.c_uint64()
        _lib.check(self.lib.sc_batch_result(self.handle, int(i), _lib.as_u32p(polys), _lib.as_u32p(chal),
                                            ctypes.byref(cs)))
        return polys, chal[: nv - 1], int(cs.value)

    def info(self) -> dict:
        nb, st, tl = ctypes.c_uint64(), ctypes.c_int32(), ctypes.c_uint64()
        _lib.check(self.lib.sc_batch_info(self.handle, ctypes.byref(nb), ctypes.byref(st), ctypes.byref(tl)))
        return {"device_bytes": nb.value, "stages": st.value, "tiles": tl.value}
