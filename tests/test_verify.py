"""The verifier (zkduel_hip.verify) on transcripts produced by the oracle: it accepts honest
proofs, final-point values included, and rejects each kind of tampering. CPU only."""

from __future__ import annotations

import random

import numpy as np
import pytest

from oracle import sumcheck_ref as R
from zkduel_hip import Composition, get_field, get_spec
from zkduel_hip.verify import VerificationError, interpolate, transcript_challenge, verify


def _mle(table, point, p):
    """Multilinear extension at point; variable k is bit k of the index (pairs 2i, 2i+1 fold first)."""
    total = 0
    for i, x in enumerate(table):
        w = 1
        for k, r in enumerate(point):
            w = w * (r if (i >> k) & 1 else 1 - r) % p
        total = (total + x * w) % p
    return total


def _proof(bw, comp, nv, mode, seed=3):
    f = get_field(bw)
    p, L = f.modulus, f.limbs
    rng = random.Random(seed)
    tables = [[rng.randrange(p) for _ in range(1 << nv)] for _ in range(comp.num_tables)]
    terms = list(comp.terms())
    chal = R.sha3_challenger(p, L) if mode == "sha3" else R.fixed_challenger(1)
    claim, rounds, rs = R.prove(tables, terms, p, nv, chal)
    r_final = chal(nv - 1, rounds[-1]) % p
    point = rs + [r_final]
    finals = [_mle(t, point, p) for t in tables]
    enc = lambda xs: f.ints_to_limbs([f.montgomery_encode(x) for x in xs])
    polys = np.stack([enc(g) for g in rounds])
    return f, claim, polys, enc(rs) if rs else np.zeros((0, L), np.uint32), enc([r_final])[0], enc(finals)


@pytest.mark.parametrize("bw", [32, 64, 128, 256])
@pytest.mark.parametrize("mode", ["sha3", "fixed"])
def test_verifier_accepts_honest_and_rejects_tampered(bw, mode):
    comp = Composition.from_spec(get_spec("poly_abc_plus_de"))
    f, claim, polys, chal, rf, finals = _proof(bw, comp, 5, mode)
    verify(bw, comp, polys, chal, challenge_mode=mode, claim=claim, r_final=rf, final_values=finals)
    bad = polys.copy()
    bad[2, 1, 0] ^= 1
    with pytest.raises(VerificationError):
        verify(bw, comp, bad, chal, challenge_mode=mode)
    with pytest.raises(VerificationError):
        verify(bw, comp, polys, chal, challenge_mode=mode, claim=claim + 1)
    bad_f = finals.copy()
    bad_f[0, 0] ^= 1
    with pytest.raises(VerificationError):
        verify(bw, comp, polys, chal, challenge_mode=mode, r_final=rf, final_values=bad_f)
    if mode == "sha3":
        bad_c = chal.copy()
        bad_c[1, 0] ^= 1
        with pytest.raises(VerificationError):
            verify(bw, comp, polys, bad_c, challenge_mode=mode)


def test_transcript_and_interpolation_match_the_oracle():
    f = get_field(256)
    p = f.modulus
    rng = random.Random(9)
    g = [rng.randrange(p) for _ in range(5)]
    limbs = f.ints_to_limbs([f.montgomery_encode(x) for x in g])
    assert transcript_challenge(f, 7, limbs) == R.transcript_challenge(7, g, p, f.limbs)
    r = rng.randrange(p)
    assert interpolate(g, r, p) == R.lagrange_eval(g, r, p)
    assert [interpolate(g, t, p) for t in range(5)] == g
