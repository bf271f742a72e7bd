"""SumCheck verifier for proofs produced by the library: a cheap guard a server or caller can
run on every result.

Given the round polynomials g_j (values at t = 0..d, Montgomery limbs), the challenges, and
optionally the final-point step of the proof (``Proof.final()``), it checks:

* every challenge is the one the transcript rule derives from its round polynomial
  (SHA3-256 of "zkduel:field-sweep:v1" || u32le(round) || u32le(limbs of g_round), read
  little-endian mod p; cuda/src/field_sumcheck.cu:857-907), or 1 for fixed challenges;
* g_j(0) + g_j(1) = g_{j-1}(r_{j-1}) for every round j >= 1 (and = claim for j = 0, when a
  claim is given), with g_{j-1}(r) interpolated from its values at 0..d;
* g_last(r_final) = C(f_1(r), ..., f_T(r)) for the final table values.

Host-side Python integers; it verifies, it does not prove.
"""

from __future__ import annotations

import hashlib
from typing import Sequence

import numpy as np

from .fields import FieldSpec, get_field
from .prover import Composition

TRANSCRIPT_TAG = b"zkduel:field-sweep:v1"


class VerificationError(ValueError):
    """A proof failed one of the verifier's checks (the message names which)."""


def transcript_challenge(field: FieldSpec, round_index: int, g_limbs: np.ndarray) -> int:
    """Canonical challenge for one round from the Montgomery limbs of its polynomial."""
    msg = bytearray(TRANSCRIPT_TAG)
    msg += int(round_index).to_bytes(4, "little")
    for w in np.asarray(g_limbs, dtype=np.uint32).reshape(-1):
        msg += int(w).to_bytes(4, "little")
    digest = hashlib.sha3_256(bytes(msg)).digest()
    return int.from_bytes(digest, "little") % field.modulus


def interpolate(values: Sequence[int], r: int, p: int) -> int:
    """g(r) for the polynomial of degree < len(values) with g(t) = values[t], t = 0..d."""
    n = len(values)
    total = 0
    for t in range(n):
        num, den = 1, 1
        for s in range(n):
            if s != t:
                num = num * (r - s) % p
                den = den * (t - s) % p
        total = (total + values[t] * num * pow(den, -1, p)) % p
    return total


def evaluate_composition(comp: Composition, point: Sequence[int], p: int) -> int:
    total = 0
    for coeff, idx in comp.terms():
        prod = coeff % p
        for v in idx:
            prod = prod * point[v] % p
        total = (total + prod) % p
    return total


def verify(bit_width: int, comp: Composition, round_polys: np.ndarray, challenges: np.ndarray, *,
           challenge_mode: str = "sha3", claim: int | None = None, r_final: np.ndarray | None = None,
           final_values: np.ndarray | None = None) -> None:
    """Raises VerificationError naming the first failed check; returns None if all pass.

    round_polys [num_vars][d+1][L], challenges [num_vars-1][L], r_final [L] and
    final_values [T][L] are Montgomery limbs as the library returns them."""
    f = get_field(bit_width)
    p = f.modulus
    polys = np.asarray(round_polys, dtype=np.uint32)
    nv = polys.shape[0]
    g = [[f.montgomery_decode(x) for x in f.limbs_to_ints(polys[j])] for j in range(nv)]
    chal = np.asarray(challenges, dtype=np.uint32).reshape(-1, f.limbs)
    r = [f.montgomery_decode(x) for x in f.limbs_to_ints(chal[: nv - 1])]
    if len(r) != nv - 1:
        raise VerificationError(f"expected {nv - 1} challenges, got {len(r)}")

    def expected_challenge(j: int) -> int:
        if challenge_mode == "sha3":
            return transcript_challenge(f, j, polys[j])
        if challenge_mode == "fixed":
            return 1
        raise ValueError("challenge_mode must be 'sha3' or 'fixed' for transcript checks")

    if challenge_mode != "explicit":
        for j in range(nv - 1):
            if r[j] != expected_challenge(j):
                raise VerificationError(f"challenge {j} does not follow the transcript")
    if claim is not None and (g[0][0] + g[0][1]) % p != claim % p:
        raise VerificationError("g_0(0) + g_0(1) differs from the claim")
    for j in range(1, nv):
        if (g[j][0] + g[j][1]) % p != interpolate(g[j - 1], r[j - 1], p):
            raise VerificationError(f"round {j}: g_j(0) + g_j(1) != g_(j-1)(r_(j-1))")
    if final_values is not None:
        if r_final is None:
            raise ValueError("final_values needs r_final")
        rf = f.montgomery_decode(f.limbs_to_ints(np.asarray(r_final, dtype=np.uint32).reshape(1, -1))[0])
        if challenge_mode != "explicit" and rf != expected_challenge(nv - 1):
            raise VerificationError("final challenge does not follow the transcript")
        vals = [f.montgomery_decode(x) for x in f.limbs_to_ints(np.asarray(final_values, dtype=np.uint32))]
        if len(vals) != comp.num_tables:
            raise VerificationError(f"expected {comp.num_tables} final values, got {len(vals)}")
        if interpolate(g[nv - 1], rf, p) != evaluate_composition(comp, vals, p):
            raise VerificationError("g_last(r_final) != C(final table values)")
