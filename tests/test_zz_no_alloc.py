"""sc_prove_noalloc (the entry point that allocates nothing) against sc_prove, the entry point
it is a variant of: bit-identical round polynomials, challenges and checksum, on two calls in
a row whose second input is smaller, at every width; and the size limit and mode errors.

The test allocates its own device buffers (input tables and outputs) through the HIP runtime;
the entry point under test allocates nothing.
"""

from __future__ import annotations

import ctypes
import sys

import numpy as np
import pytest

from zkduel_hip import Composition, Context, Proof, _lib, get_field, get_spec

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def hip():
    rt = ctypes.CDLL("libamdhip64.so")
    rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    rt.hipFree.argtypes = [ctypes.c_void_p]
    rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return rt


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


class DeviceBuf:
    def __init__(self, rt, nbytes):
        self.rt = rt
        self.ptr = ctypes.c_void_p()
        assert rt.hipMalloc(ctypes.byref(self.ptr), max(int(nbytes), 16)) == 0
        self.nbytes = int(nbytes)

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        assert self.rt.hipMemcpy(self.ptr, arr.ctypes.data, arr.nbytes, H2D) == 0

    def download(self, shape, dtype):
        out = np.zeros(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        assert self.rt.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H) == 0
        return out

    def free(self):
        if self.ptr:
            self.rt.hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()


def prove_both(ctx, hip, bw, comp, nv, mode, seed):
    """(sc_prove result, sc_prove_noalloc result) on the same device tables."""
    f = get_field(bw)
    L, npts = f.limbs, comp.degree + 1
    with Proof(ctx, bw, comp, nv, input="uniform", seed=seed) as src:
        host_tables = src.tables()
    tables = DeviceBuf(hip, host_tables.nbytes)
    polys_d = DeviceBuf(hip, nv * npts * L * 4)
    chal_d = DeviceBuf(hip, max(nv - 1, 1) * L * 4)
    cs_d = DeviceBuf(hip, 8)
    try:
        tables.upload(host_tables)
        cs, keep = comp.bind(f)
        # the existing entry point, on the same device tables (borrowed input)
        inp = _lib.ScInput(_lib.SC_INPUT_DEVICE_MONT, 0, tables.ptr.value, 0, 0)
        want_polys = np.zeros((nv, npts, L), dtype=np.uint32)
        want_chal = np.zeros((max(nv - 1, 1), L), dtype=np.uint32)
        want_cs = ctypes.c_uint64()
        _lib.check(ctx.lib.sc_prove(ctx.handle, bw, ctypes.byref(cs), nv, ctypes.byref(inp), mode,
                                    None, _lib.as_u32p(want_polys), _lib.as_u32p(want_chal),
                                    ctypes.byref(want_cs)))
        _lib.check(ctx.lib.sc_prove_noalloc(ctx.handle, bw, ctypes.byref(cs), nv, tables.ptr, mode,
                                            polys_d.ptr, chal_d.ptr, cs_d.ptr))
        ctx.sync()
        got_polys = polys_d.download((nv, npts, L), np.uint32)
        got_chal = chal_d.download((max(nv - 1, 1), L), np.uint32)
        got_cs = int(cs_d.download((1,), np.uint64)[0])
        del keep
        return ((want_polys, want_chal[: nv - 1], int(want_cs.value)),
                (got_polys, got_chal[: nv - 1], got_cs))
    finally:
        for b in (tables, polys_d, chal_d, cs_d):
            b.free()


# (width, first call num_vars, second call num_vars): the first call runs pass, point-parallel
# (255-bit) and tail rounds; the second is smaller, so stale scratch contents would show
CASES = [(32, 20, 9), (64, 20, 11), (128, 18, 7), (256, 18, 10)]


@pytest.mark.parametrize("bw,nv1,nv2", CASES)
@pytest.mark.parametrize("mode", ["sha3", "fixed"])
def test_noalloc_matches_sc_prove_two_calls(ctx, hip, bw, nv1, nv2, mode):
    m = _lib.CHALLENGE_MODES[mode]
    big = Composition.from_spec(get_spec("poly_abc_plus_de"))
    small = Composition.from_terms(4, [(3, [0, 1, 1]), (-1, [2]), (1, [3, 0])])
    for comp, nv, seed in ((big, nv1, 7), (small, nv2, 8)):
        say(f"noalloc: bw={bw} num_vars={nv} tables={comp.num_tables} mode={mode}")
        want, got = prove_both(ctx, hip, bw, comp, nv, m, seed)
        assert np.array_equal(got[0], want[0]), (bw, nv, "round polynomials")
        assert np.array_equal(got[1], want[1]), (bw, nv, "challenges")
        assert got[2] == want[2], (bw, nv, "checksum")


def test_noalloc_rejects_oversized_and_explicit(ctx):
    f = get_field(256)
    comp = Composition.from_spec(get_spec("poly_abc_plus_de"))
    cs, keep = comp.bind(f)
    dummy = ctypes.c_void_p(16)  # never dereferenced: both calls fail their checks on the host
    # 0.75 x 5 tables x 2^21 entries x 32 bytes = 240 MiB > the 128 MiB area
    rc = ctx.lib.sc_prove_noalloc(ctx.handle, 256, ctypes.byref(cs), 21, dummy, _lib.SC_CHALLENGE_SHA3,
                                  dummy, None, None)
    assert rc == _lib.SC_ERR_UNSUPPORTED
    assert b"exceed" in ctx.lib.sc_last_error()
    rc = ctx.lib.sc_prove_noalloc(ctx.handle, 256, ctypes.byref(cs), 8, dummy, _lib.SC_CHALLENGE_EXPLICIT,
                                  dummy, None, None)
    assert rc == _lib.SC_ERR_UNSUPPORTED
    del keep
