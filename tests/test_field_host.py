"""CPU check of the field layer: the device templates of csrc/field.hpp built for the host
(tests/field_host_harness.cpp) against Python integers. The GPU tests repeat the same checks
on the device; this one runs without a GPU, on every change."""

from __future__ import annotations

import ctypes
import random
import subprocess
from pathlib import Path

import numpy as np
import pytest

from zkduel_hip.fields import get_field

ROOT = Path(__file__).resolve().parents[1]
SRC = ROOT / "tests" / "field_host_harness.cpp"
INC = ROOT / "sumcheck-triton_amd" / "csrc"


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    out = tmp_path_factory.mktemp("fieldhost") / "field_host.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", f"-I{INC}", str(SRC), "-o", str(out)],
                   check=True)
    lib = ctypes.CDLL(str(out))
    u32p = ctypes.POINTER(ctypes.c_uint32)
    lib.field_host_op.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint64, u32p, u32p, u32p]
    lib.field_host_op.restype = None

    def op(bw, code, a, b):
        f = get_field(bw)
        A, B = f.ints_to_limbs(a), f.ints_to_limbs(b)
        out = np.zeros_like(A)
        lib.field_host_op(bw, code, len(a), A.ctypes.data_as(u32p), B.ctypes.data_as(u32p), out.ctypes.data_as(u32p))
        return f.limbs_to_ints(out)

    return op


def operands(bw, n=4000):
    f = get_field(bw)
    p = f.modulus
    rng = random.Random(bw * 3)
    edge = [0, 1, 2, 3, p - 3, p - 2, p - 1, f.r_mod, p - f.r_mod, (1 << (32 * f.limbs - 1)) % p,
            (1 << 32) % p, ((1 << 32) - 1) % p, (1 << 64) % p, p >> 1, (p + 1) >> 1]
    a = [x for x in edge for _ in edge] + [rng.randrange(p) for _ in range(n)]
    b = [y for _ in edge for y in edge] + [rng.randrange(p) for _ in range(n)]
    # operands whose words are all-ones or zero stress every carry and borrow path
    for _ in range(n):
        a.append(sum(rng.choice((0, 0xFFFFFFFF, 1, 0x80000000)) << (32 * i) for i in range(f.limbs)) % p)
        b.append(p - 1 - rng.randrange(1 << 16))
    return a, b


@pytest.mark.parametrize("bw", [32, 64, 128, 256])
def test_host_field_ops(harness, bw):
    f = get_field(bw)
    p = f.modulus
    a, b = operands(bw)
    rinv = f.r_inv
    want_mul = [x * y * rinv % p for x, y in zip(a, b)]
    assert harness(bw, 0, a, b) == want_mul
    assert harness(bw, 3, a, b) == want_mul  # generic CIOS agrees with the specialised product
    assert harness(bw, 4, a, b) == want_mul  # product scanning (host form of the device MAC)
    assert harness(bw, 1, a, b) == [(x + y) % p for x, y in zip(a, b)]
    assert harness(bw, 2, a, b) == [(x - y) % p for x, y in zip(a, b)]
