"""The four prime fields and their Montgomery conventions (src/zkduel/fields.py:230-304).

The device library hard-codes these moduli. The reference guards its *truncated* CIOS
routine with a modulus bound (src/zkduel/fields.py:118-227); the HIP kernels keep CIOS's
full carry word, so they are exact for every odd modulus below 2^(32L) and need no such guard.
``FieldSpec`` still validates what the library requires (odd, fits in L limbs, correct
nprime) and refuses anything else.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

MASK32 = 0xFFFFFFFF


class UnsupportedFieldError(ValueError):
    """The modulus is not one the library was compiled for, or is malformed."""


def montgomery_nprime(modulus: int) -> int:
    """-modulus^-1 mod 2^32."""
    return (-pow(int(modulus), -1, 1 << 32)) & MASK32


def _unsupported_reason(modulus: int, limbs: int, nprime: int) -> str | None:
    """Why the library cannot run Montgomery arithmetic with these parameters, or None.

    The device kernels need an odd modulus of at most 32 * limbs bits and the nprime the
    Montgomery product uses; any such modulus is exact (they keep the full carry word)."""
    bits = 32 * limbs
    if limbs <= 0:
        return "the limb count must be positive"
    if modulus <= 2 or not modulus & 1:
        return "Montgomery form needs an odd modulus above 2"
    if modulus.bit_length() > bits:
        return f"the modulus has {modulus.bit_length()} bits, more than {bits}"
    want = montgomery_nprime(modulus)
    if nprime != want:
        return f"nprime is {nprime:#x}; -p^-1 mod 2^32 is {want:#x}"
    return None


@dataclass(frozen=True)
class FieldSpec:
    name: str
    bit_width: int
    limbs: int
    modulus: int
    nprime: int

    def __post_init__(self) -> None:
        reason = _unsupported_reason(int(self.modulus), int(self.limbs), int(self.nprime))
        if reason is not None:
            raise UnsupportedFieldError(f"field {self.name!r} (p = {int(self.modulus):#x}): {reason}")

    @property
    def r_mod(self) -> int:
        return (1 << (32 * self.limbs)) % self.modulus

    @property
    def r_inv(self) -> int:
        return pow(self.r_mod, -1, self.modulus)

    def to_limbs(self, value: int) -> tuple[int, ...]:
        x = int(value) % self.modulus
        return tuple((x >> (32 * i)) & MASK32 for i in range(self.limbs))

    def montgomery_encode(self, value: int) -> int:
        return (int(value) % self.modulus) * self.r_mod % self.modulus

    def montgomery_decode(self, value: int) -> int:
        return int(value) * self.r_inv % self.modulus

    def montgomery_limbs(self, value: int) -> tuple[int, ...]:
        return self.to_limbs(self.montgomery_encode(value))

    # Vectorised conversions between Python ints and [n][L] uint32 limb arrays.
    def ints_to_limbs(self, values) -> np.ndarray:
        vals = [int(v) % self.modulus for v in values]
        out = np.empty((len(vals), self.limbs), dtype=np.uint32)
        for i in range(self.limbs):
            out[:, i] = [(v >> (32 * i)) & MASK32 for v in vals]
        return out

    def limbs_to_ints(self, limbs: np.ndarray) -> list[int]:
        arr = np.asarray(limbs, dtype=np.uint64).reshape(-1, self.limbs)
        vals = [0] * arr.shape[0]
        for i in range(self.limbs - 1, -1, -1):
            col = arr[:, i].tolist()
            vals = [(v << 32) | int(w) for v, w in zip(vals, col)]
        return vals


FIELD_SPECS: dict[int, FieldSpec] = {
    32: FieldSpec("p32_2^32_minus_5", 32, 1, 0xFFFFFFFB, 0xCCCCCCCD),
    64: FieldSpec("p64_goldilocks", 64, 2, 0xFFFFFFFF00000001, 0xFFFFFFFF),
    128: FieldSpec("p128_mersenne_2^127_minus_1", 128, 4, (1 << 127) - 1, 0x00000001),
    256: FieldSpec(
        "p255_bls12_381_scalar",
        256,
        8,
        int("73eda753299d7d483339d80809a1d80553bda402fffe5bfeffffffff00000001", 16),
        0xFFFFFFFF,
    ),
}


def get_field(bit_width: int) -> FieldSpec:
    """The field of the given width (32, 64, 128 or 256); ValueError for any other."""
    spec = FIELD_SPECS.get(int(bit_width))
    if spec is None:
        widths = "/".join(map(str, FIELD_SPECS))
        raise ValueError(f"no field of bit width {bit_width}: the library has {widths}")
    return spec


def field_for_modulus(q: int) -> FieldSpec:
    for spec in FIELD_SPECS.values():
        if spec.modulus == int(q):
            return spec
    raise UnsupportedFieldError(
        f"q={q:#x} is not one of the compiled fields (2^32-5, Goldilocks, 2^127-1, BLS12-381 Fr)"
    )
