"""Composition expressions, with the reference package's rules (src/zkduel/_expr.py:27-110).

A term is ``["a", "b"]`` (coefficient 1) or ``(coeff, ["a", "b"])``; variables are ordered
by first appearance; the degree is the longest term. ``encode_expression_with_coeffs`` gives
the flat encoding the C ABI takes.
"""

from __future__ import annotations

from typing import Any, Sequence

Expression = Sequence[Any]
NormalizedExpression = tuple[tuple[int, tuple[str, ...]], ...]


def _is_coeff_term(term: Any) -> bool:
    if not isinstance(term, (list, tuple)) or len(term) != 2:
        return False
    coeff, factors = term
    if isinstance(coeff, str) or isinstance(factors, str):
        return False
    try:
        int(coeff)
    except (TypeError, ValueError):
        return False
    try:
        return all(isinstance(v, str) for v in factors)
    except TypeError:
        return False


def normalize_expression(expression: Expression) -> NormalizedExpression:
    terms: list[tuple[int, tuple[str, ...]]] = []
    for term in expression:
        if _is_coeff_term(term):
            coeff, factors = term
            terms.append((int(coeff), tuple(factors)))
            continue
        if isinstance(term, str):
            raise TypeError("Expression terms must be sequences of variable names")
        try:
            factors = tuple(term)
        except TypeError as exc:
            raise TypeError("Expression terms must be sequences") from exc
        if not all(isinstance(v, str) for v in factors):
            raise TypeError(
                "Expression terms must contain variable names, or be "
                "(coefficient, variable_names) pairs"
            )
        terms.append((1, factors))
    if not terms:
        raise ValueError("expression must contain at least one term")
    return tuple(terms)


def expression_vars(expression: Expression) -> list[str]:
    seen: list[str] = []
    for _c, term in normalize_expression(expression):
        for v in term:
            if v not in seen:
                seen.append(v)
    return seen


def expression_degree(expression: Expression) -> int:
    return max(len(term) for _c, term in normalize_expression(expression))


def encode_expression_with_coeffs(
    expression: Expression,
) -> tuple[list[str], tuple[int, ...], tuple[int, ...], tuple[int, ...], tuple[int, ...]]:
    """``(var_names, coeffs, term_lens, term_offsets, term_vars)``."""
    terms = normalize_expression(expression)
    names = expression_vars(expression)
    idx = {v: i for i, v in enumerate(names)}
    coeffs, lens, offs, flat = [], [], [], []
    for c, term in terms:
        offs.append(len(flat))
        coeffs.append(c)
        lens.append(len(term))
        flat.extend(idx[v] for v in term)
    return names, tuple(coeffs), tuple(lens), tuple(offs), tuple(flat)


def encode_expression(expression: Expression):
    """``(var_names, term_lens, term_offsets, term_vars)`` (src/zkduel/_expr.py:70-84)."""
    names, _c, lens, offs, flat = encode_expression_with_coeffs(expression)
    return names, lens, offs, flat
