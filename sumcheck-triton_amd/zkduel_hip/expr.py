"""Composition expressions as the reference package accepts them (src/zkduel/_expr.py:27-110).

The rules, which this module implements and ``tests/test_host.py`` pins:

* An expression is a non-empty iterable of terms (empty: ``ValueError``).
* A term is a *weighted term* when it is a list or tuple of exactly two items ``(c, names)``
  where ``c`` is not a string and ``int(c)`` succeeds (so Python ints, bools, numpy integers
  and floats, truncated), and ``names`` is not a string and iterates to strings only. Its
  coefficient is ``int(c)``.
* Any other term must be an iterable of variable names (strings) and has coefficient 1. A bare
  string, a non-iterable, or an iterable holding a non-string is a ``TypeError``.
* Variables are numbered in order of first appearance over the terms; the degree is the length
  of the longest term; a variable repeated in a term is a power.

``encode_expression_with_coeffs`` flattens the result into the arrays of the C ABI's
``sc_composition`` (the reference's WorkloadSpec encoding, workload_specs.py:6-17).
"""

from __future__ import annotations

from typing import Any, Sequence

Expression = Sequence[Any]
Term = tuple[int, tuple[str, ...]]
NormalizedExpression = tuple[Term, ...]


def _names_or_none(obj: Any) -> tuple[str, ...] | None:
    """obj as a tuple of variable names, or None if it is a string, not iterable, or holds a
    non-string."""
    if isinstance(obj, str):
        return None
    try:
        items = tuple(obj)
    except TypeError:
        return None
    for item in items:
        if not isinstance(item, str):
            return None
    return items


def _int_or_none(obj: Any) -> int | None:
    """int(obj) for anything int() converts, except strings; else None."""
    if isinstance(obj, str):
        return None
    try:
        return int(obj)
    except (TypeError, ValueError):
        return None


def _weighted(term: Any) -> Term | None:
    """(coefficient, names) when term is a weighted term, else None."""
    if not isinstance(term, (list, tuple)) or len(term) != 2:
        return None
    coeff = _int_or_none(term[0])
    names = _names_or_none(term[1])
    if coeff is None or names is None:
        return None
    return coeff, names


def _term(term: Any) -> Term:
    weighted = _weighted(term)
    if weighted is not None:
        return weighted
    if isinstance(term, str):
        raise TypeError(f"term {term!r} is a string: write a term as a sequence of variable names")
    try:
        factors = tuple(term)
    except TypeError as exc:
        raise TypeError(f"term {term!r} is not a sequence of variable names") from exc
    if any(not isinstance(v, str) for v in factors):
        raise TypeError(f"term {term!r} must hold variable names, or be a (coefficient, names) pair")
    return 1, factors


def normalize_expression(expression: Expression) -> NormalizedExpression:
    """Every term as ``(int coefficient, tuple of names)``, in the given order."""
    terms = tuple(_term(t) for t in expression)
    if len(terms) == 0:
        raise ValueError("an expression needs at least one term")
    return terms


def _first_appearance(terms: NormalizedExpression) -> list[str]:
    order: dict[str, None] = {}
    for _coeff, names in terms:
        for name in names:
            order.setdefault(name, None)
    return list(order)


def expression_vars(expression: Expression) -> list[str]:
    """Variable names in order of first appearance."""
    return _first_appearance(normalize_expression(expression))


def expression_degree(expression: Expression) -> int:
    """Length of the longest term."""
    return max(len(names) for _coeff, names in normalize_expression(expression))


def encode_expression_with_coeffs(
    expression: Expression,
) -> tuple[list[str], tuple[int, ...], tuple[int, ...], tuple[int, ...], tuple[int, ...]]:
    """``(var_names, coeffs, term_lens, term_offsets, term_vars)``: term k's factors are
    ``term_vars[term_offsets[k] : term_offsets[k] + term_lens[k]]``, indices into var_names."""
    terms = normalize_expression(expression)
    names = _first_appearance(terms)
    index = {name: i for i, name in enumerate(names)}
    lens = tuple(len(factors) for _c, factors in terms)
    offsets = tuple(sum(lens[:k]) for k in range(len(lens)))
    flat = tuple(index[name] for _c, factors in terms for name in factors)
    return names, tuple(c for c, _f in terms), lens, offsets, flat


def encode_expression(expression: Expression):
    """``(var_names, term_lens, term_offsets, term_vars)`` (the reference's _expr.py:70-84)."""
    names, _coeffs, lens, offsets, flat = encode_expression_with_coeffs(expression)
    return names, lens, offsets, flat
