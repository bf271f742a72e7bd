"""ctypes binding of the oracle's C twin (liboracle.so) — TEST INFRASTRUCTURE ONLY."""

from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

from . import sumcheck_ref as R

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
_lib = None

u32p = ctypes.POINTER(ctypes.c_uint32)
i32p = ctypes.POINTER(ctypes.c_int32)


def load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            subprocess.run(["make", "-C", str(HERE)], check=True, capture_output=True)
        lib = ctypes.CDLL(str(LIB))
        lib.oracle_prove.restype = ctypes.c_int
        lib.oracle_prove.argtypes = [
            ctypes.c_int, ctypes.c_int, ctypes.c_int, i32p, i32p, i32p, u32p, ctypes.c_int,
            ctypes.c_int, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64, u32p, ctypes.c_int,
            u32p, ctypes.c_int, u32p, u32p, u32p,
        ]
        lib.oracle_claim_final.restype = ctypes.c_int
        lib.oracle_claim_final.argtypes = [
            ctypes.c_int, ctypes.c_int, ctypes.c_int, i32p, i32p, i32p, u32p, ctypes.c_int,
            ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64, u32p, ctypes.c_int, u32p, u32p,
        ]
        lib.oracle_sha3_256.restype = None
        lib.oracle_sha3_256.argtypes = [ctypes.c_char_p, ctypes.c_uint64, ctypes.c_char_p]
        lib.oracle_mont_mul.restype = None
        lib.oracle_mont_mul.argtypes = [ctypes.c_int, ctypes.c_uint64, u32p, u32p, u32p]
        _lib = lib
    return _lib


def _limbs(vals, L):
    out = np.zeros((max(len(vals), 1), L), dtype=np.uint32)
    for i, v in enumerate(vals):
        out[i] = R.to_limbs(int(v), L)
    return out


def prove(bit_width, num_tables, terms, num_vars, *, input_kind="uniform", seed=1, instance=0,
          selector_mask=0, host_tables=None, challenge_mode="sha3", challenges=None, threads=None):
    """Returns (polys_mont [nv][d+1][L] uint32, r_mont [nv-1][L], final [T][2][L])."""
    lib = load()
    p, L = R.field(bit_width)
    lens = np.array([len(idx) for _c, idx in terms], dtype=np.int32)
    offs = np.array(np.cumsum([0] + [len(idx) for _c, idx in terms])[:-1], dtype=np.int32)
    flat = np.array([i for _c, idx in terms for i in idx], dtype=np.int32)
    coeffs = _limbs([c % p for c, _idx in terms], L)
    deg = int(lens.max())
    kinds = {"affine": 0, "uniform": 1, "host": 2}
    modes = {"fixed": 0, "sha3": 1, "explicit": 2}
    host = None
    if input_kind == "host":
        host = np.ascontiguousarray(host_tables, dtype=np.uint32)
    expl = _limbs([c % p for c in (challenges or [])], L)
    polys = np.zeros((num_vars, deg + 1, L), dtype=np.uint32)
    rr = np.zeros((max(num_vars - 1, 1), L), dtype=np.uint32)
    final = np.zeros((num_tables, 2, L), dtype=np.uint32)
    threads = threads or min(8, os.cpu_count() or 1)
    rc = lib.oracle_prove(
        bit_width, num_tables, len(terms), lens.ctypes.data_as(i32p), offs.ctypes.data_as(i32p),
        flat.ctypes.data_as(i32p), coeffs.ctypes.data_as(u32p), num_vars, kinds[input_kind],
        seed, instance, selector_mask, host.ctypes.data_as(u32p) if host is not None else None,
        modes[challenge_mode], expl.ctypes.data_as(u32p), threads, polys.ctypes.data_as(u32p),
        rr.ctypes.data_as(u32p), final.ctypes.data_as(u32p))
    if rc != 0:
        raise RuntimeError(f"oracle_prove failed ({rc})")
    return polys, rr[: num_vars - 1], final


def claim_and_final(bit_width, num_tables, terms, num_vars, point_mont, *, seed=1, instance=0,
                    selector_mask=0, threads=None):
    """For the uniform tables keyed by (seed, instance): (claim [L], final [T][L]) in Montgomery
    limbs: the sum of the composition over the hypercube, and every table's multilinear
    extension at point_mont ([num_vars][L] Montgomery, variable 0 = the first round's)."""
    lib = load()
    p, L = R.field(bit_width)
    lens = np.array([len(idx) for _c, idx in terms], dtype=np.int32)
    offs = np.array(np.cumsum([0] + [len(idx) for _c, idx in terms])[:-1], dtype=np.int32)
    flat = np.array([i for _c, idx in terms for i in idx], dtype=np.int32)
    coeffs = _limbs([c % p for c, _idx in terms], L)
    point = np.ascontiguousarray(point_mont, dtype=np.uint32).reshape(num_vars, L)
    claim = np.zeros(L, dtype=np.uint32)
    final = np.zeros((num_tables, L), dtype=np.uint32)
    threads = threads or min(16, os.cpu_count() or 1)
    rc = lib.oracle_claim_final(
        bit_width, num_tables, len(terms), lens.ctypes.data_as(i32p), offs.ctypes.data_as(i32p),
        flat.ctypes.data_as(i32p), coeffs.ctypes.data_as(u32p), num_vars, seed, instance, selector_mask,
        point.ctypes.data_as(u32p), threads, claim.ctypes.data_as(u32p), final.ctypes.data_as(u32p))
    if rc != 0:
        raise RuntimeError(f"oracle_claim_final failed ({rc})")
    return claim, final


def sha3_256(data: bytes) -> bytes:
    out = ctypes.create_string_buffer(32)
    load().oracle_sha3_256(data, len(data), out)
    return out.raw


def mont_mul(bit_width, a, b):
    p, L = R.field(bit_width)
    A, B = _limbs(a, L), _limbs(b, L)
    out = np.zeros_like(A)
    load().oracle_mont_mul(bit_width, len(a), A.ctypes.data_as(u32p), B.ctypes.data_as(u32p), out.ctypes.data_as(u32p))
    return [R.from_limbs(row) for row in out[: len(a)]]
