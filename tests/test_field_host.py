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
    lib.field_host_work.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint64, u32p, u32p, u32p, u32p]
    lib.field_host_work.restype = None
    lib.r29_host.argtypes = [ctypes.c_int, ctypes.c_uint64, u32p, u32p]
    lib.r29_host.restype = None
    lib.uniform_host.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, u32p]
    lib.uniform_host.restype = None
    u64p = ctypes.POINTER(ctypes.c_uint64)
    lib.gl_host_mul_sum.argtypes = [ctypes.c_uint64, u64p, u64p, u64p, u64p, u64p]
    lib.gl_host_mul_sum.restype = None
    lib.r29_fold_tab_host.argtypes = [ctypes.c_int, ctypes.c_uint64, u32p, u32p, u32p, u32p]
    lib.r29_fold_tab_host.restype = None
    lib.word_columns_host.argtypes = [ctypes.c_int, ctypes.c_uint64, u64p, u32p]
    lib.word_columns_host.restype = None
    lib.interleave_host.argtypes = [ctypes.c_uint64, u64p, u32p, u32p, u64p]
    lib.interleave_host.restype = None
    lib.gl_host_accum.argtypes = [ctypes.c_uint64, ctypes.POINTER(ctypes.c_int32), u64p, u64p]
    lib.gl_host_accum.restype = ctypes.c_uint64
    lib.gl_host_work_accum.argtypes = [ctypes.c_uint64, ctypes.POINTER(ctypes.c_int32), u64p, u64p]
    lib.gl_host_work_accum.restype = ctypes.c_uint64
    lib.gl_host_fast.argtypes = [ctypes.c_int, ctypes.c_uint64, u64p, u64p, u64p]
    lib.gl_host_fast.restype = None

    def op(bw, code, a, b):
        f = get_field(bw)
        A, B = f.ints_to_limbs(a), f.ints_to_limbs(b)
        out = np.zeros_like(A)
        lib.field_host_op(bw, code, len(a), A.ctypes.data_as(u32p), B.ctypes.data_as(u32p), out.ctypes.data_as(u32p))
        return f.limbs_to_ints(out)

    def work(bw, code, a, b, c):
        f = get_field(bw)
        A, B, C = f.ints_to_limbs(a), f.ints_to_limbs(b), f.ints_to_limbs(c)
        out = np.zeros_like(A)
        lib.field_host_work(bw, code, len(a), A.ctypes.data_as(u32p), B.ctypes.data_as(u32p),
                            C.ctypes.data_as(u32p), out.ctypes.data_as(u32p))
        return f.limbs_to_ints(out)

    def r29(code, arr):
        arr = np.ascontiguousarray(arr, dtype=np.uint32)
        out = np.zeros_like(arr)
        lib.r29_host(code, arr.shape[0], arr.ctypes.data_as(u32p), out.ctypes.data_as(u32p))
        return out

    def gl_mul_sum(a1, b1, a2, b2):
        arrs = [np.ascontiguousarray(x, dtype=np.uint64) for x in (a1, b1, a2, b2)]
        out = np.zeros_like(arrs[0])
        lib.gl_host_mul_sum(len(out), *[x.ctypes.data_as(u64p) for x in arrs], out.ctypes.data_as(u64p))
        return [int(v) for v in out]

    op.work = work
    op.r29 = r29
    def interleave(words):
        w = np.ascontiguousarray(words, dtype=np.uint64)
        e, o = np.zeros(len(w), np.uint32), np.zeros(len(w), np.uint32)
        back = np.zeros_like(w)
        lib.interleave_host(len(w), w.ctypes.data_as(u64p), e.ctypes.data_as(u32p), o.ctypes.data_as(u32p),
                            back.ctypes.data_as(u64p))
        return e, o, back

    def gl_accum(ops, a, b, work_form=False):
        o = np.ascontiguousarray(ops, dtype=np.int32)
        A, B = (np.ascontiguousarray(x, dtype=np.uint64) for x in (a, b))
        fn = lib.gl_host_work_accum if work_form else lib.gl_host_accum
        return int(fn(len(o), o.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                      A.ctypes.data_as(u64p), B.ctypes.data_as(u64p)))

    def gl_fast(code, a, b):
        A, B = (np.ascontiguousarray(x, dtype=np.uint64) for x in (a, b))
        out = np.zeros_like(A)
        lib.gl_host_fast(code, len(A), A.ctypes.data_as(u64p), B.ctypes.data_as(u64p), out.ctypes.data_as(u64p))
        return [int(v) for v in out]

    op.gl_fast = gl_fast

    def fold_tab(mont, lo, hi, rm):
        lo, hi = (np.ascontiguousarray(x, dtype=np.uint32) for x in (lo, hi))
        rm = np.ascontiguousarray(rm, dtype=np.uint32)
        out = np.zeros((lo.shape[0], 9), dtype=np.uint32)
        lib.r29_fold_tab_host(mont, lo.shape[0], lo.ctypes.data_as(u32p), hi.ctypes.data_as(u32p),
                              rm.ctypes.data_as(u32p), out.ctypes.data_as(u32p))
        return out

    op.fold_tab = fold_tab

    def word_columns(bw, cols):
        L = get_field(bw).limbs
        c = np.ascontiguousarray(cols, dtype=np.uint64).reshape(-1, L)
        out = np.zeros(c.shape, dtype=np.uint32)
        lib.word_columns_host(bw, c.shape[0], c.ctypes.data_as(u64p), out.ctypes.data_as(u32p))
        return [sum(int(x) << (32 * k) for k, x in enumerate(row)) for row in out.tolist()]

    op.word_columns = word_columns

    op.gl_mul_sum = gl_mul_sum
    op.gl_accum = gl_accum
    op.interleave = interleave
    op.lib = lib
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


@pytest.mark.parametrize("bw", [32, 64, 128, 256])
def test_host_working_form(harness, bw):
    """The kernels' working-form path (Montgomery in, working arithmetic, Montgomery out)
    against Python integers, on edge and random operands; for the 255-bit field this is the
    radix-2^29 representation with lazily reduced values."""
    f = get_field(bw)
    p, rinv = f.modulus, f.r_inv
    a, b = operands(bw, n=1500)
    rng = random.Random(bw + 11)
    c = [rng.choice((0, 1, p - 1, rng.randrange(p))) for _ in a]
    mont = lambda x, y: x * y * rinv % p
    assert harness.work(bw, 0, a, b, c) == [mont(x, y) for x, y in zip(a, b)]
    assert harness.work(bw, 1, a, b, c) == [(x + mont(z, y - x)) % p for x, y, z in zip(a, b, c)]
    x7 = [(x + 7 * (y - x)) % p for x, y in zip(a, b)]
    assert harness.work(bw, 2, a, b, c) == [(-mont(mont(mont(t, t), t), z)) % p for t, z in zip(x7, c)]
    assert harness.work(bw, 3, a, b, c) == [39 * mont(x, y) % p for x, y in zip(a, b)]
    want4 = []
    for x, y, z in zip(a, b, c):
        xs = sum((x + t * (y - x)) for t in range(8))
        want4.append((8 * mont(x, y) - mont(z, xs % p)) % p)
    assert harness.work(bw, 4, a, b, c) == want4
    assert harness.work(bw, 5, a, b, c) == [(mont(mont(x, t), t) + mont(z, x)) % p for x, t, z in zip(a, x7, c)]
    # -x(t) alone in a fresh accumulator (a term like `- g` of a degree-7 composition): x(7)
    # of the 255-bit working form comes to 29p, so the subtraction's padding must exceed it
    want6 = [(-(x + (i % 8) * (y - x))) % p for i, (x, y) in enumerate(zip(a, b))]
    assert harness.work(bw, 6, a, b, c) == want6
    assert harness.work(bw, 7, a, b, c) == want6
    # round 1 folding straight from the caller's Montgomery words, and round 0's scaled read
    assert harness.work(bw, 8, a, b, c) == [(x + mont(z, y - x)) % p for x, y, z in zip(a, b, c)]
    inv32 = pow(32, -1, p) if bw == 256 else 1
    assert harness.work(bw, 9, a, b, c) == [x * inv32 % p for x in a]
    # products whose second operands are interpolated values left unnormalized
    x2s = [(x + 2 * (y - x)) % p for x, y in zip(a, b)]
    x3s = [(x + 3 * (y - x)) % p for x, y in zip(a, b)]
    # op 11: a(2) = 2b - a, b(2) = 2a - c, times b (Montgomery products)
    assert harness.work(bw, 11, a, b, c) == [mont(mont((2 * y - x) % p, (2 * x - z) % p), y)
                                             for x, y, z in zip(a, b, c)]
    assert harness.work(bw, 10, a, b, c) == [(mont(x, u3) + mont(x, u2) + mont(z, u2)) % p
                                             for x, z, u2, u3 in zip(a, c, x2s, x3s)]


def test_host_goldilocks_product_sum_one_reduction(harness):
    """Goldilocks a1 b1 + a2 b2 with a single reduction, on any 64-bit operands (the kernels
    pass it values left unreduced below 2^64): the result is below 2^64 and congruent to
    (a1 b1 + a2 b2) 2^-64 mod p. Operands near 2^64 make the 129-bit sum carry out."""
    f = get_field(64)
    p, rinv = f.modulus, f.r_inv
    rng = random.Random(64064)
    edge = [0, 1, p - 1, p, p + 1, (1 << 64) - 1, (1 << 64) - 2, (1 << 63), (1 << 32) - 1, 1 << 32]
    cols = [[], [], [], []]
    for x in edge:
        for y in edge:
            for c, v in zip(cols, (x, y, y, x)):
                c.append(v)
    for _ in range(20000):
        for c in cols:
            c.append(rng.choice((rng.randrange(1 << 64), (1 << 64) - 1 - rng.randrange(1 << 20),
                                 rng.randrange(p))))
    got = harness.gl_mul_sum(*cols)
    for g, a1, b1, a2, b2 in zip(got, *cols):
        assert g < (1 << 64) and g % p == (a1 * b1 + a2 * b2) * rinv % p, (a1, b1, a2, b2)


def test_host_r29_pack_and_reduce(harness):
    """Radix-2^29 helpers: unpack/pack round trip of 256-bit words, and the top-limb reduction
    (result congruent mod p, below p + 2^241, limbs normalized) on limbs up to 2^30 with a top
    limb up to 2^32."""
    p = get_field(256).modulus
    rng = np.random.default_rng(5)
    w = rng.integers(0, 1 << 32, size=(4000, 8), dtype=np.uint64).astype(np.uint32)
    w[:8] = 0xFFFFFFFF
    w[8:16] = 0
    assert np.array_equal(harness.r29(0, w), w)
    limbs = rng.integers(0, 1 << 30, size=(4000, 9), dtype=np.uint64)
    limbs[:, 8] = rng.integers(0, 1 << 32, size=4000, dtype=np.uint64)
    limbs[:4] = [(1 << 30) - 1] * 8 + [(1 << 32) - 1]
    out = harness.r29(1, limbs.astype(np.uint32))
    for row_in, row_out in zip(limbs.tolist(), out.tolist()):
        v = sum(int(x) << (29 * i) for i, x in enumerate(row_in))
        r = sum(int(x) << (29 * i) for i, x in enumerate(row_out))
        assert r % p == v % p and r < p + (1 << 241) and all(x < (1 << 29) for x in row_out)


def _words(x):
    return [(x >> (32 * k)) & 0xFFFFFFFF for k in range(8)]


@pytest.mark.parametrize("bw", [32, 64, 128, 256])
def test_host_uniform_generator_matches_oracle(harness, bw):
    """The device input generator (prng.hpp, built for the host) gives the oracle's uniform
    entries: at 255 bits, where 9.4% of attempts are rejected, across indices with one, two
    and three attempts."""
    from oracle import sumcheck_ref as R

    p, L = R.field(bw)
    for seed, inst, table in ((5, 3, 2), (41, 0, 11), (1, 7, 0)):
        key = R.stream_key(seed, inst, table)
        n = 3000
        out = np.zeros(n * L, dtype=np.uint32)
        harness.lib.uniform_host(bw, key, 1000, n, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
        got = [sum(int(w) << (32 * k) for k, w in enumerate(out[i * L:(i + 1) * L])) for i in range(n)]
        assert got == [R.uniform_element(key, 1000 + i, p, L) for i in range(n)], (bw, seed)


def test_host_r29_scale_small(harness):
    """The 255-bit multiply by a small integer c (|c| <= 7), used for selector factors whose
    round-0 values are 0 or 1: on normalized inputs up to the largest the kernels pass (top
    limb below 2^29: products, coefficients, interpolated values up to 29p), the result is
    congruent to c x, at most 3p (3p itself for c < 0 and x = 0), with every limb below 2^29."""
    p = get_field(256).modulus
    rng = random.Random(12)
    xs = [0, 1, p - 1, p, 13 * p - 1, 29 * p + (1 << 244), (1 << 261) - 1]
    xs += [rng.randrange(1 << 261) for _ in range(400)] + [rng.randrange(13 * p) for _ in range(400)]
    rows = []
    for x in xs:
        for c in range(-7, 8):
            rows.append([(x >> (29 * i)) & ((1 << 29) - 1) for i in range(9)] + [c & 0xFFFFFFFF])
    out = harness.r29(2, np.array(rows, dtype=np.uint64).astype(np.uint32))
    for row, res in zip(rows, out.tolist()):
        x = sum(int(v) << (29 * i) for i, v in enumerate(row[:9]))
        c = row[9] - (1 << 32) if row[9] >= 1 << 31 else row[9]
        v = sum(int(w) << (29 * i) for i, w in enumerate(res[:9]))
        assert all(w < (1 << 29) for w in res[:9]), (x, c)
        assert v <= 3 * p, (x, c)
        assert v % p == c * x % p, (x, c)


def test_host_r29_fold_by_table(harness):
    """The 255-bit fold by the challenge's table (no Montgomery reduction; the quotient from
    the top columns in double precision) on stored values across their whole range
    [0, p + 2^241), edge cases included, and on the caller's Montgomery input: the result is
    congruent to lo + r (hi - lo) in the working form (x 2^261), below p + 2^241, with every limb
    below 2^29. Also values placed next to multiples of p, where a wrong quotient would show."""
    p = get_field(256).modulus
    W = 1 << 261
    top = p + (1 << 241)
    rng = random.Random(11)
    rs = [1, 2, p - 1, (p + 1) // 2, 1 << 200] + [rng.randrange(p) for _ in range(6)]
    for r in rs:
        rm = _words(r * (1 << 256) % p)
        lo = [0, 0, top - 1, top - 1, p, p - 1, 1] + [rng.randrange(top) for _ in range(1500)]
        hi = [top - 1, 0, 0, top - 1, 0, p, top - 2] + [rng.randrange(top) for _ in range(1500)]  # odd count: the last one takes fold1_tab
        # hi within 2 of lo: D = hi - lo + 3p close to 3p
        for _ in range(200):
            x = rng.randrange(p)
            y = (x + rng.choice([-2, -1, 0, 1, 2])) % p
            lo.append(x)
            hi.append(y)
        out = harness.fold_tab(0, [_words(x) for x in lo], [_words(x) for x in hi], rm)
        for a, b, row in zip(lo, hi, out.tolist()):
            v = sum(int(x) << (29 * i) for i, x in enumerate(row))
            assert all(x < (1 << 29) for x in row), (r, a, b)
            assert v < top, (r, a, b)
            assert v % p == (a + r * (b - a)) % p, (r, a, b)
        # straight from Montgomery input X = x 2^256 (canonical): working value x 2^261 = 32 X
        mlo = [0, p - 1, p - 1] + [rng.randrange(p) for _ in range(600)]
        mhi = [p - 1, 0, p - 1] + [rng.randrange(p) for _ in range(600)]
        if len(mlo) % 2 == 0:  # an odd count: the last entry goes through the one-fold form
            mlo.append(1)
            mhi.append(2)
        out = harness.fold_tab(1, [_words(x) for x in mlo], [_words(x) for x in mhi], rm)
        for a, b, row in zip(mlo, mhi, out.tolist()):
            v = sum(int(x) << (29 * i) for i, x in enumerate(row))
            assert all(x < (1 << 29) for x in row) and v < top, (r, a, b)
            assert v % p == 32 * (a + r * (b - a)) % p, (r, a, b)
    assert W > top


@pytest.mark.parametrize("bw", [32, 64, 128, 256])
def test_host_word_column_sums(harness, bw):
    """Sums of canonical values added word by word into 64-bit columns (no carries), then
    reduced once (reduce_word_columns, the block-partial sums of the 8-limb finalize): equal to
    the sum mod p for 1 to 4096 values, all-(p - 1) and all-zero inputs, exact multiples of p,
    and column sums whose carries reach the top word."""
    f = get_field(bw)
    p, L = f.modulus, f.limbs
    rng = random.Random(bw + 7)
    cases = []
    for n in [1, 2, 3, 255, 256, 4096]:
        cases.append([rng.randrange(p) for _ in range(n)])
    cases += [[p - 1] * 4096, [0] * 17, [p - 1] * 3, [1] * 5]
    # sums that land on or next to a multiple of p, and just past one by less than p 2^-20 (the
    # quotient is then one short and V - q p exceeds p: past 2^(32L) at 32 and 64 bits)
    for k in [1, 2, 1000, 4095]:
        xs = [rng.randrange(p) for _ in range(k)]
        s = sum(xs)
        m = ((s // p) + 1) * p
        for target in (m, m - 1, m + 1, m + 5, m + (1 << 33), m + (p >> 21)):
            extra = target - s
            if 0 <= extra < p:
                cases.append(xs + [extra])
    cols = []
    for xs in cases:
        cols.append([sum((x >> (32 * k)) & 0xFFFFFFFF for x in xs) for k in range(L)])
    got = harness.word_columns(bw, cols)
    for xs, g in zip(cases, got):
        assert g == sum(xs) % p, len(xs)


def test_host_keccak_bit_interleaving(harness):
    """The wave Keccak holds each 64-bit word as its even bits and its odd bits: check both
    halves bit by bit against the definition, the join back, and that a 64-bit rotation by r
    is the two 32-bit rotations the kernel uses (r even: both halves by r/2; r odd: the even
    half takes the odd half rotated by (r+1)/2, the odd half the even half by (r-1)/2)."""
    rng = random.Random(25)
    words = [0, 1, 1 << 63, (1 << 64) - 1, 0x5555555555555555, 0xAAAAAAAAAAAAAAAA] + \
        [rng.getrandbits(64) for _ in range(2000)]
    e, o, back = harness.interleave(words)
    for w, ev, od, bk in zip(words, e.tolist(), o.tolist(), back.tolist()):
        assert ev == sum(((w >> (2 * i)) & 1) << i for i in range(32))
        assert od == sum(((w >> (2 * i + 1)) & 1) << i for i in range(32))
        assert bk == w

    def rotl(x, r, bits):
        r %= bits
        return ((x << r) | (x >> (bits - r))) & ((1 << bits) - 1) if r else x

    for r in range(64):
        rotated = [rotl(w, r, 64) for w in words[:200]]
        re, ro, _ = harness.interleave(rotated)
        ev, od = e[:200].tolist(), o[:200].tolist()
        if r % 2 == 0:
            want_e, want_o = [rotl(x, r // 2, 32) for x in ev], [rotl(x, r // 2, 32) for x in od]
        else:
            want_e, want_o = [rotl(x, (r + 1) // 2, 32) for x in od], [rotl(x, (r - 1) // 2, 32) for x in ev]
        assert re.tolist() == want_e and ro.tolist() == want_o, r


@pytest.mark.parametrize("work_form", [False, True], ids=["accum", "work_columns"])
def test_host_goldilocks_wide_accumulator(harness, work_form):
    """The Goldilocks running sums hold an integer W whose value is W R^-1: values enter as
    X 2^64, products A B raw, and one reduction at the end (Accum: W = top 2^128 + w_hi 2^64 +
    w_lo; the working form's Acc: four 64-bit column sums of 32-bit words). Checked against
    Python integers on long runs of near-2^64 operands, where every carry out of bit 63 and
    of bit 127 happens many times."""
    p = (1 << 64) - (1 << 32) + 1
    rinv = pow(1 << 64, -1, p)
    rng = random.Random(64)
    for case in range(6):
        n = 3000 if case < 5 else 20000
        hi_bias = case % 2 == 0
        ops = [rng.choice((0, 1, 2, 2)) for _ in range(n)]
        if case == 4:
            ops = [2] * n  # products only: bit 127 carries on every few adds
        draw = (lambda: (1 << 64) - 1 - rng.randrange(1 << 20)) if hi_bias else (lambda: rng.getrandbits(64))
        a = [draw() for _ in range(n)]
        b = [draw() for _ in range(n)]
        want = 0
        for o, x, y in zip(ops, a, b):
            if o == 0:
                want += x
            elif o == 1:
                want -= x
            else:
                want += x * y * rinv
        assert harness.gl_accum(ops, a, b, work_form) == want % p, case


def test_host_goldilocks_working_primitives(harness):
    """The Goldilocks working form's carry-free product and its reduction, on raw 64-bit
    words (the kernels pass products and lazy sums anywhere below 2^64): the 128-bit product
    is exact, x 2^-64 is congruent and below 2^64; x + d and x - d for any x and d < p are
    congruent and below 2^64 (the difference canonical when x < p); canon is x mod p."""
    p = (1 << 64) - (1 << 32) + 1
    rinv = pow(1 << 64, -1, p)
    rng = random.Random(6464)
    M = (1 << 64) - 1
    edge = [0, 1, 2, (1 << 32) - 1, 1 << 32, (1 << 32) + 1, p - 2, p - 1, p, p + 1, M - 1, M,
            1 << 63, (1 << 63) - 1, 0xFFFFFFFF00000000, 0x00000000FFFFFFFF]
    a = [x for x in edge for _ in edge]
    b = [y for _ in edge for y in edge]
    for _ in range(30000):
        a.append(rng.choice((rng.getrandbits(64), M - rng.randrange(1 << 24), rng.randrange(p),
                             rng.randrange(1 << 32) << 32, rng.randrange(1 << 33))))
        b.append(rng.choice((rng.getrandbits(64), M - rng.randrange(1 << 24), rng.randrange(p))))
    lo = harness.gl_fast(5, a, b)
    hi = harness.gl_fast(4, a, b)
    red = harness.gl_fast(0, a, b)
    for x, y, l, h, r in zip(a, b, lo, hi, red):
        assert (h << 64) | l == x * y, (x, y)
        assert r <= M and r % p == x * y * rinv % p, (x, y)
    bc = [y % p for y in b]  # d and y canonical
    for x, d, s in zip(a, bc, harness.gl_fast(1, a, bc)):
        assert s <= M and s % p == (x + d) % p, (x, d)
    for x, d, s in zip(a, bc, harness.gl_fast(2, a, bc)):
        assert s <= M and s % p == (x - d) % p, (x, d)
        if x < p:
            assert s < p, (x, d)
    for x, c in zip(a, harness.gl_fast(3, a, b)):
        assert c == x % p, x
