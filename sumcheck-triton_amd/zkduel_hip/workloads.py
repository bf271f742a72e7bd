"""The reference's 32 fixed compositions, with its two naming schemes.

Two tables of the reference are restated here as data, because goldens recorded by the
reference depend on them exactly:

* the kernel-facing specs (``benchmarks/field_sweep/workload_specs.py:20-749``): name, id,
  polynomial and the **table numbering** the CUDA/Triton kernels use (tables are numbered
  by first appearance in the polynomial, except ``zk_complete_add_1``, whose kernels number
  ``lambda`` after ``x_p``);
* the expression registry and its aliases (``src/zkduel/workloads.py:11-263``), used by the
  package API ``get_workload`` / ``sumcheck``.

Each polynomial is parsed once into ``WorkloadSpec`` (coeffs, term_lens, term_offsets,
term_vars), the same encoding the C ABI takes.
"""

from __future__ import annotations

import re
from dataclasses import dataclass

from .expr import Expression

# Deterministic placeholder challenge of the zkPHIRE presets (src/zkduel/workloads.py:9).
alpha = 7


@dataclass(frozen=True)
class WorkloadSpec:
    name: str
    polynomial: str
    ident: int
    var_names: tuple[str, ...]  # table numbering
    coeffs: tuple[int, ...]
    term_lens: tuple[int, ...]
    term_offsets: tuple[int, ...]
    term_vars: tuple[int, ...]

    @property
    def num_vars(self) -> int:  # number of tables (the reference's field name)
        return len(self.var_names)

    @property
    def degree(self) -> int:
        return max(self.term_lens)

    @property
    def terms(self) -> int:
        return len(self.term_lens)

    def expression(self) -> Expression:
        """The registry-style expression: ``[(coeff, [names...]), ...]``."""
        out = []
        for c, n, o in zip(self.coeffs, self.term_lens, self.term_offsets):
            names = [self.var_names[i] for i in self.term_vars[o : o + n]]
            out.append(names if c == 1 else (c, names))
        return out


# (id, spec name, registry name, polynomial[, explicit table order])
_TABLE: tuple[tuple, ...] = (
    (1, "poly_a", "poly_a", "a"),
    (2, "poly_ab", "poly_ab", "a*b"),
    (3, "poly_ab_plus_c", "poly_ab_plus_c", "a*b + c"),
    (4, "poly_abc", "poly_abc", "a*b*c"),
    (5, "poly_aabbc", "poly_aabbc", "a*a*b*b*c"),
    (6, "poly_abc_plus_de", "poly_abc_plus_de", "a*b*c + d*e"),
    (7, "poly_abcg_plus_deg", "poly_abcg_plus_deg", "a*b*c*g + d*e*g"),
    (8, "zk_verifiable_asics", "verifiable_asics", "qadd*a + qadd*b + qmul*a*b"),
    (9, "zk_spartan_1", "spartan_1", "A*B*fz - C*fz"),
    (10, "zk_spartan_2", "spartan_2", "ABC*Z"),
    (11, "zk_witness_non_id", "witness_non_id_point",
     "q_non-id-point*y*y - q_non-id-point*x*x*x - 5*q_non-id-point"),
    (12, "zk_complete_add_1", "complete_addition_1",
     "q_add*x_q*x_q*lambda - 2*q_add*x_q*x_p*lambda + q_add*x_p*x_p*lambda - q_add*x_q*y_q"
     " + q_add*x_q*y_p + q_add*x_p*y_q - q_add*x_p*y_p",
     ("q_add", "x_q", "x_p", "lambda", "y_q", "y_p")),
    (13, "zk_complete_add_7", "complete_addition_7",
     "q_add*x_r - q_add*x_q - q_add*x_p*x_r*beta + q_add*x_p*x_q*beta"),
    (14, "zk_complete_add_8", "complete_addition_8",
     "q_add*y_r - q_add*y_q - q_add*x_p*y_r*beta + q_add*x_p*y_q*beta"),
    (15, "zk_witness_id_point_1", "witness_id_point_1",
     "q_point*x*y*y - q_point*x*x*x*x - 5*q_point*x"),
    (16, "zk_witness_id_point_2", "witness_id_point_2",
     "q_point*y*y*y - q_point*y*x*x*x - 5*q_point*y"),
    (17, "zk_incomplete_add_1", "incomplete_addition_1",
     "q_add-incomplete*x_r*x_p*x_p - 2*q_add-incomplete*x_r*x_p*x_q + q_add-incomplete*x_r*x_q*x_q"
     " + q_add-incomplete*x_q*x_p*x_p - 2*q_add-incomplete*x_q*x_q*x_p"
     " + q_add-incomplete*x_q*x_q*x_q + q_add-incomplete*x_p*x_p*x_p"
     " - 2*q_add-incomplete*x_p*x_p*x_q + q_add-incomplete*x_p*x_q*x_q"
     " - q_add-incomplete*y_p*y_p + 2*q_add-incomplete*y_p*y_q - q_add-incomplete*y_q*y_q"),
    (18, "zk_incomplete_add_2", "incomplete_addition_2",
     "q_add-incomplete*y_r*x_p - q_add-incomplete*y_r*x_q + q_add-incomplete*y_q*x_p"
     " - q_add-incomplete*y_q*x_q - q_add-incomplete*y_p*x_q + q_add-incomplete*y_p*x_r"
     " - q_add-incomplete*y_q*x_r + q_add-incomplete*y_q*x_q"),
    (19, "zk_complete_add_2", "complete_addition_2",
     "2*q_add*y_p*lambda - 3*q_add*x_p*x_p - 2*q_add*x_q*y_p*lambda*alpha"
     " + 3*q_add*x_q*x_p*x_p*alpha + 2*q_add*x_p*y_p*lambda*alpha - 3*q_add*x_p*x_p*x_p*alpha"),
    (20, "zk_complete_add_3", "complete_addition_3",
     "q_add*x_p*x_q*x_q*lambda*lambda - q_add*x_p*x_q*x_q*x_q - q_add*x_p*x_q*x_q*x_r"
     " - q_add*x_p*x_p*x_q*x_q - q_add*x_p*x_p*x_q*lambda*lambda + q_add*x_p*x_p*x_p*x_q"
     " + q_add*x_p*x_p*x_q*x_r + q_add*x_p*x_p*x_q*x_q"),
    (21, "zk_complete_add_4", "complete_addition_4",
     "q_add*x_p*x_p*x_q*x_q*lambda - q_add*x_p*x_q*x_q*x_r*lambda - q_add*x_p*x_q*x_q*y_p"
     " - q_add*x_p*x_q*x_q*y_r - q_add*x_p*x_p*x_p*x_q*lambda + q_add*x_p*x_p*x_q*x_r*lambda"
     " + q_add*x_p*x_p*x_q*y_p + q_add*x_p*x_p*x_q*y_r"),
    (22, "zk_complete_add_5", "complete_addition_5",
     "q_add*x_p*x_q*y_q*lambda*lambda + q_add*x_p*x_q*y_p*lambda*lambda - q_add*x_p*x_p*x_q*y_q"
     " - q_add*x_p*x_p*x_q*y_p - q_add*x_p*x_q*x_q*y_q - q_add*x_p*x_q*x_q*y_p"
     " - q_add*x_p*x_q*y_q*x_r - q_add*x_p*x_q*y_p*x_r"),
    (23, "zk_complete_add_6", "complete_addition_6",
     "q_add*x_p*x_p*x_q*y_q*lambda + q_add*x_p*x_p*x_q*y_p*lambda - q_add*x_p*x_q*y_p*x_r*lambda"
     " - q_add*x_p*x_q*y_q*x_r*lambda - q_add*x_p*x_q*y_p*y_p - q_add*x_p*x_q*y_p*y_q"
     " - q_add*x_p*x_q*y_p*y_r - q_add*x_p*x_q*y_q*y_r"),
    (24, "zk_complete_add_9", "complete_addition_9",
     "q_add*x_r - q_add*x_p - q_add*x_q*x_r*gamma + q_add*x_p*x_q*gamma"),
    (25, "zk_complete_add_10", "complete_addition_10",
     "q_add*y_r - q_add*y_p - q_add*x_q*y_r*gamma + q_add*x_q*y_p*gamma"),
    (26, "zk_complete_add_11", "complete_addition_11",
     "q_add*x_r - q_add*x_q*x_r*alpha + q_add*x_p*x_r*alpha - q_add*y_q*x_r*delta"
     " - q_add*y_p*x_r*delta"),
    (27, "zk_complete_add_12", "complete_addition_12",
     "q_add*y_r - q_add*x_q*y_r*alpha + q_add*x_p*y_r*alpha - q_add*y_q*y_r*delta"
     " - q_add*y_p*y_r*delta"),
    (28, "zk_vanilla_zerocheck_hp", "vanilla_zerocheck_hp",
     "qL*w1*fz1 + qR*w2*fz1 + qM*w1*w2*fz1 - qO*w3*fz1 + qC*fz1"),
    (29, "zk_vanilla_permcheck_hp", "vanilla_permcheck_hp",
     "pi*fz2 - p1*p2*fz2 + 7*phi*D1*D2*D3*fz2 - 7*N1*N2*N3*fz2"),
    (30, "zk_jellyfish_zerocheck_hp", "jellyfish_zerocheck_hp",
     "qL*w1*fz + qR*w2*fz + qM*w1*w2*fz - qO*w3*fz + q1*w1*fz + q2*w2*fz + q3*w3*fz + q4*w4*fz"
     " - qO*w5*fz + qM1*w1*w2*fz + qM2*w3*w4*fz + qH1*w1*w1*w1*w1*w1*fz"
     " + qH2*w2*w2*w2*w2*w2*fz + qH3*w3*w3*w3*w3*w3*fz + qH4*w4*w4*w4*w4*w4*fz"
     " + qECC*w1*w2*w3*w4*fz + qc*fz"),
    (31, "zk_jellyfish_permcheck_hp", "jellyfish_permcheck_hp",
     "pi*fz - p1*p2*fz + 7*phi*D1*D2*D3*D4*D5*fz - 7*N1*N2*N3*N4*N5*fz"),
    (32, "zk_opencheck", "opencheck", "y1*k1 + y2*k2 + y3*k3 + y4*k4 + y5*k5 + y6*k6"),
)

_TERM_SPLIT = re.compile(r"\s+([+-])\s+")


def parse_polynomial(polynomial: str, var_order: tuple[str, ...] | None = None) -> WorkloadSpec:
    """Parse ``"2*a*b - c + ..."`` (terms joined by `` + ``/`` - ``, factors by ``*``; an
    all-digit factor multiplies the coefficient). Tables are numbered by first appearance
    unless ``var_order`` is given."""
    parts = _TERM_SPLIT.split(" + " + polynomial.strip())
    # parts = ["", sign, term, sign, term, ...]
    terms: list[tuple[int, list[str]]] = []
    for sign, body in zip(parts[1::2], parts[2::2]):
        coeff = -1 if sign == "-" else 1
        names: list[str] = []
        for factor in body.split("*"):
            factor = factor.strip()
            if not factor:
                raise ValueError(f"empty factor in {polynomial!r}")
            if factor.isdigit():
                coeff *= int(factor)
            else:
                names.append(factor)
        if not names:
            raise ValueError(f"term without a table in {polynomial!r}")
        terms.append((coeff, names))
    order: list[str] = []
    for _c, names in terms:
        for v in names:
            if v not in order:
                order.append(v)
    if var_order is not None:
        if sorted(var_order) != sorted(order):
            raise ValueError("var_order must list exactly the polynomial's tables")
        order = list(var_order)
    idx = {v: i for i, v in enumerate(order)}
    coeffs, lens, offs, flat = [], [], [], []
    for c, names in terms:
        offs.append(len(flat))
        coeffs.append(c)
        lens.append(len(names))
        flat.extend(idx[v] for v in names)
    return WorkloadSpec("", polynomial, 0, tuple(order), tuple(coeffs), tuple(lens), tuple(offs), tuple(flat))


def _build() -> tuple[dict[str, WorkloadSpec], dict[str, str]]:
    specs: dict[str, WorkloadSpec] = {}
    registry_names: dict[str, str] = {}
    for row in _TABLE:
        ident, name, reg_name, poly = row[:4]
        order = row[4] if len(row) > 4 else None
        s = parse_polynomial(poly, order)
        specs[name] = WorkloadSpec(name, poly, ident, s.var_names, s.coeffs, s.term_lens,
                                   s.term_offsets, s.term_vars)
        registry_names[reg_name] = name
    return specs, registry_names


SPECS, _REGISTRY_TO_SPEC = _build()
SPEC_NAMES: list[str] = [row[1] for row in _TABLE]
POLY_SPEC_NAMES = SPEC_NAMES[:7]
ZK_SPEC_NAMES = SPEC_NAMES[7:]

# Alias tables of src/zkduel/workloads.py:212-255 and
# benchmarks/field_sweep/workload_specs.py:711-737.
_ZKPHIRE = [
    "verifiable_asics", "spartan_1", "spartan_2", "witness_non_id_point", "witness_id_point_1",
    "witness_id_point_2", "incomplete_addition_1", "incomplete_addition_2",
    *[f"complete_addition_{i}" for i in range(1, 13)],
    "vanilla_zerocheck_hp", "vanilla_permcheck_hp", "jellyfish_zerocheck_hp",
    "jellyfish_permcheck_hp", "opencheck",
]
ALIASES: dict[str, str] = {}
for _i, _reg in enumerate(_ZKPHIRE):
    ALIASES[f"zkphire_{_i}"] = _REGISTRY_TO_SPEC[_reg]
for _reg, _spec in _REGISTRY_TO_SPEC.items():
    if _reg != _spec:
        ALIASES[_reg] = _spec
for _short, _spec in [("a", "poly_a"), ("ab", "poly_ab"), ("ab_plus_c", "poly_ab_plus_c"),
                      ("abc", "poly_abc"), ("aabbc", "poly_aabbc"),
                      ("abc_plus_de", "poly_abc_plus_de"), ("abcg_plus_deg", "poly_abcg_plus_deg"),
                      ("linear", "poly_a"), ("quadratic", "poly_ab"),
                      ("cubic_plus_quad", "poly_abc_plus_de"),
                      ("hyperplonk_zero", "zk_vanilla_zerocheck_hp"),
                      ("hyperplonk_perm", "zk_vanilla_permcheck_hp"),
                      ("hyperplonk_open", "zk_opencheck")]:
    ALIASES[_short] = _spec


def canonical_workload_name(name: str) -> str:
    return ALIASES.get(name, name)


def get_spec(name: str) -> WorkloadSpec:
    try:
        return SPECS[canonical_workload_name(name)]
    except KeyError as exc:
        known = ", ".join(sorted([*SPECS, *ALIASES]))
        raise KeyError(f"unknown workload {name!r}; known: {known}") from exc


def get_workload(name: str) -> Expression:
    """Expression for a registry, spec or alias name (src/zkduel/workloads.py:258-263)."""
    return get_spec(name).expression()


WORKLOADS: dict[str, Expression] = {}
for _reg, _spec in _REGISTRY_TO_SPEC.items():
    WORKLOADS[_reg] = SPECS[_spec].expression()
for _alias in ALIASES:
    WORKLOADS[_alias] = get_workload(_alias)
