"""CPU tests of the host layer: registry, expressions, fields, and the C ABI library's exports.

No GPU is touched: the library is loaded and its symbols are checked, and creating a context
without a device must raise (the product has no CPU fallback).
"""

from __future__ import annotations

import re
import subprocess
from pathlib import Path

import pytest

from zkduel_hip import _lib
from zkduel_hip.expr import encode_expression_with_coeffs, expression_degree, expression_vars, normalize_expression
from zkduel_hip.fields import FIELD_SPECS, FieldSpec, UnsupportedFieldError, get_field, montgomery_nprime
from zkduel_hip.workloads import ALIASES, SPEC_NAMES, SPECS, WORKLOADS, get_spec, get_workload, parse_polynomial

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "sumcheck_hip.h"

# (name, tables, degree, terms) of cuda/src/field_sumcheck.cu:106-139
REFERENCE_SHAPES = {
    "poly_a": (1, 1, 1), "poly_ab": (2, 2, 1), "poly_ab_plus_c": (3, 2, 2), "poly_abc": (3, 3, 1),
    "poly_aabbc": (3, 5, 1), "poly_abc_plus_de": (5, 3, 2), "poly_abcg_plus_deg": (6, 4, 2),
    "zk_verifiable_asics": (4, 3, 3), "zk_spartan_1": (4, 3, 2), "zk_spartan_2": (2, 2, 1),
    "zk_witness_non_id": (3, 4, 3), "zk_complete_add_1": (6, 4, 7), "zk_complete_add_7": (5, 4, 4),
    "zk_complete_add_8": (5, 4, 4), "zk_witness_id_point_1": (3, 5, 3),
    "zk_witness_id_point_2": (3, 5, 3), "zk_incomplete_add_1": (6, 4, 12),
    "zk_incomplete_add_2": (7, 3, 8), "zk_complete_add_2": (6, 5, 6), "zk_complete_add_3": (5, 6, 8),
    "zk_complete_add_4": (7, 6, 8), "zk_complete_add_5": (7, 6, 8), "zk_complete_add_6": (8, 6, 8),
    "zk_complete_add_9": (5, 4, 4), "zk_complete_add_10": (5, 4, 4), "zk_complete_add_11": (8, 4, 5),
    "zk_complete_add_12": (8, 4, 5), "zk_vanilla_zerocheck_hp": (9, 4, 5),
    "zk_vanilla_permcheck_hp": (11, 5, 4), "zk_jellyfish_zerocheck_hp": (22, 7, 17),
    "zk_jellyfish_permcheck_hp": (15, 7, 4), "zk_opencheck": (12, 2, 6),
}


def test_registry_shapes_match_reference_table():
    assert len(SPECS) == 32 and set(SPEC_NAMES) == set(REFERENCE_SHAPES)
    for name, (t, d, k) in REFERENCE_SHAPES.items():
        s = get_spec(name)
        assert (s.num_vars, s.degree, s.terms) == (t, d, k), name
    assert [SPECS[n].ident for n in SPEC_NAMES] == sorted(SPECS[n].ident for n in SPEC_NAMES)


def test_complete_add_1_uses_kernel_numbering():
    s = get_spec("zk_complete_add_1")
    assert s.var_names == ("q_add", "x_q", "x_p", "lambda", "y_q", "y_p")
    assert s.term_vars[:4] == (0, 1, 1, 3)  # q * xq * xq * lambda


def test_aliases_and_registry_names():
    assert get_spec("zkphire_8").name == "zk_complete_add_1"
    assert get_spec("zkphire_24").name == "zk_opencheck"
    assert get_spec("hyperplonk_perm").name == "zk_vanilla_permcheck_hp"
    assert get_spec("abc").name == "poly_abc"
    assert len([a for a in ALIASES if a.startswith("zkphire_")]) == 25
    assert WORKLOADS["poly_abc"] == [["a", "b", "c"]]
    expr = get_workload("vanilla_permcheck_hp")
    assert (7, ["phi", "D1", "D2", "D3", "fz2"]) in expr
    with pytest.raises(KeyError):
        get_spec("no_such_workload")


def test_parse_polynomial_handles_hyphenated_names_and_coefficients():
    s = parse_polynomial("q-x*y*y - 5*q-x + 2*y*3")
    assert s.var_names == ("q-x", "y")
    assert s.coeffs == (1, -5, 6)
    assert s.term_vars == (0, 1, 1, 0, 1)
    with pytest.raises(ValueError):
        parse_polynomial("a*b", ("a", "c"))


def test_expression_rules_follow_reference():
    expr = [["a", "b"], (3, ["c", "a"]), (-1, ("d",))]
    assert normalize_expression(expr) == ((1, ("a", "b")), (3, ("c", "a")), (-1, ("d",)))
    assert expression_vars(expr) == ["a", "b", "c", "d"]
    assert expression_degree(expr) == 2
    names, coeffs, lens, offs, flat = encode_expression_with_coeffs(expr)
    assert (coeffs, lens, offs, flat) == ((1, 3, -1), (2, 2, 1), (0, 2, 4), (0, 1, 2, 0, 3))
    with pytest.raises(ValueError):
        normalize_expression([])
    with pytest.raises(TypeError):
        normalize_expression(["ab"])


def test_expression_rule_edge_cases():
    """Each rule of the reference's _expr.py:11-54, case by case: what counts as a weighted
    term, what is a plain term of coefficient 1, and what is an error."""
    import numpy as np

    # int()-convertible coefficients: numpy integers, bools, floats (truncated)
    expr = [(np.int64(5), ["a", "b"]), (np.uint32(2), ("c",)), (True, ["a"]), (2.9, ["b"])]
    assert normalize_expression(expr) == ((5, ("a", "b")), (2, ("c",)), (1, ("a",)), (2, ("b",)))
    # a two-item term of two names is a plain product, not a weighted term
    assert normalize_expression([["x", "y"], ("y", "z")]) == ((1, ("x", "y")), (1, ("y", "z")))
    # a string coefficient or a string of names is not a weighted term: the pair is then a
    # plain term holding a non-name, so an error
    with pytest.raises(TypeError):
        normalize_expression([("3", ["a"])])
    with pytest.raises(TypeError):
        normalize_expression([(3, "ab")])
    with pytest.raises(TypeError):
        normalize_expression([(None, ["a"])])
    with pytest.raises(TypeError):
        normalize_expression([5])
    with pytest.raises(TypeError):
        normalize_expression([["a", 1, "b"]])
    # repeated variables are powers; order is first appearance across terms
    names, coeffs, lens, offs, flat = encode_expression_with_coeffs([["b", "b", "a"], (-2, ["c", "b"])])
    assert names == ["b", "a", "c"] and coeffs == (1, -2) and lens == (3, 2) and offs == (0, 3)
    assert flat == (0, 0, 1, 2, 0)
    assert expression_degree([["a"], (4, ["a", "a", "a", "b"])]) == 4
    # the oracle's restatement follows the same rules
    from oracle import sumcheck_ref as R

    for e in (expr, [["x", "y"], ("y", "z")], [["b", "b", "a"], (-2, ["c", "b"])]):
        assert R.normalize_expression(e) == normalize_expression(e)
        assert R.expression_vars(e) == expression_vars(e)
    for bad in ([("3", ["a"])], [(3, "ab")], [5], [["a", 1]]):
        with pytest.raises(TypeError):
            R.normalize_expression(bad)


def test_field_specs():
    for bw, f in FIELD_SPECS.items():
        assert f.limbs * 32 == bw
        assert f.nprime == montgomery_nprime(f.modulus)
        x = 123456789
        assert f.montgomery_decode(f.montgomery_encode(x)) == x
        assert f.limbs_to_ints(f.ints_to_limbs([0, 1, f.modulus - 1])) == [0, 1, f.modulus - 1]
    with pytest.raises(UnsupportedFieldError):
        FieldSpec("even", 64, 2, 1 << 40, 0)
    with pytest.raises(UnsupportedFieldError):
        FieldSpec("wrong nprime", 32, 1, 0xFFFFFFFB, 1)
    with pytest.raises(ValueError):
        get_field(48)


def _header_functions() -> list[str]:
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\*?(sc_\w+)\(", text, re.M)))


def test_header_declares_what_the_binding_binds():
    declared = _header_functions()
    bound = sorted(name for name, _r, _a in _lib.SIGNATURES)
    assert declared == bound


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for name in _header_functions():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (sc_\w+)$", out, re.M))
    assert set(_header_functions()) <= exported
    assert lib.sc_version().decode().startswith("sumcheck_hip")


def test_library_has_gfx950_code_object():
    # the offload bundle of our own build names its target triple
    assert b"amdgcn-amd-amdhsa--gfx950" in _lib.LIB_PATH.read_bytes()


def test_context_without_device_fails_loudly():
    import torch  # noqa: F401  (only to learn whether this host has a device)

    if torch.cuda.is_available():
        pytest.skip("a device is present")
    from zkduel_hip.prover import Context

    with pytest.raises((ValueError, RuntimeError)):
        Context(0)


def test_cli_flag_grammar_matches_reference_executable():
    from zkduel_hip.cli import parse_args

    a = parse_args(["--rounds", "3", "10", "--bit-widths", "64", "256", "--workloads", "zkphire_8",
                    "poly_abc", "--warmups", "0", "--repeats", "1", "--challenge-mode", "sha3",
                    "--point-mode", "specialized", "--out", "x.csv", "--zk-only"])
    assert a["rounds"] == [3, 10] and a["bit_widths"] == [64, 256]
    assert a["workloads"] == ["zkphire_8", "poly_abc"] and a["zk_only"]
    assert (a["warmups"], a["repeats"], a["challenge_mode"], a["point_mode"]) == (0, 1, "sha3", "specialized")
    for bad in (["--block", "3"], ["--timing-mode", "x"], ["--nope"], ["--layout", "diag"], ["--seed"]):
        with pytest.raises(SystemExit) as e:
            parse_args(bad)
        assert e.value.code == 2


def test_driver_columns_are_the_reference_csv_header():
    from zkduel_hip.driver import CSV_COLUMNS

    header = ("workload,polynomial,backend,field,arithmetic,layout,point_mode,bit_width,limbs,modulus_hex,"
              "rounds,N,vars,degree,terms,warmups,repeats,timing_mode,challenge_mode,first_ms,median_ms,"
              "eval_median_ms,challenge_median_ms,fold_median_ms,total_median_ms,input_mib,checksum,status")
    assert ",".join(CSV_COLUMNS) == header  # cuda/src/field_sumcheck.cu:1336


def test_autotuner_cache_roundtrip(tmp_path):
    """The schedule cache file round-trips (CPU only: no proof is run)."""
    import json

    from zkduel_hip import autotune

    path = tmp_path / "tune.json"
    autotune._cache.clear()
    autotune._cache["64/5/3/26"] = {"max_pass_blocks": 1024, "tail_pairs": 512, "ms": 2.7, "candidates": {}}
    autotune._store(path)
    autotune._cache.clear()
    autotune._load(path)
    assert autotune._cache["64/5/3/26"]["max_pass_blocks"] == 1024
    assert json.loads(path.read_text())["64/5/3/26"]["tail_pairs"] == 512
    autotune._cache.clear()
