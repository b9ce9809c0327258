"""Pins the oracle's update rule, boundary predicate and post-pass with
hand-derived known answers (tests/kat_cases.py), and shows that each answer
catches a plausible mistake: the oracle is rebuilt with one planted mistake
at a time (ORACLE_MUTANT, oracle/propag_oracle.c) and every case that names
that mistake in its `catches` must then fail.

Planted mistakes and the cases that catch them:
  1 ct kept despite a boundary on the line          fallback_taken
  2 line-of-sight walk started at the cell          line_clear
  3 boundary on the faster side (> for <)           step_field
  4 azimuth difference with wrap-around             azimuth_no_wrap
  5 neighbours in i-outer / j-inner order           tie_order
  6 ties keep the last candidate (<= for <)         tie_order
  7 a neighbour referencing the cell not skipped    self_reference
  8 a neighbour at exactly max_time spreads         max_time_gate
  9 post-pass clears times equal to max_time        max_time_gate, post_equal_kept
"""
import os

import pytest

import kat_cases
import oracle

CASES = kat_cases.all_cases()
MUTANTS = sorted({m for c in CASES for m in c["catches"]})


def run_case(case):
    g = oracle.grid(case["W"], case["H"], 1.0, 1.0, float(case["max_time"]))
    res = oracle.propagate(g, case["speed"], case["az"], case["ecc"], case["time"],
                           case["refs_x"], case["refs_y"], max_sweeps=case["sweeps"])
    return kat_cases.mismatches(case, res["time"], res["refs_x"], res["refs_y"], res["boundaries"])


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_known_answer_holds(oracle_lib, case):
    assert run_case(case) == []


def test_walk_geometry_of_the_line_cases(oracle_lib):
    """The DDA paths the two line cases rely on (the DDA itself is pinned by
    the reference's line_to known answers in test_oracle_propag.py)."""
    assert oracle.dda(3, 1, 0, 0) == [(3, 1), (2, 1), (1, 1), (0, 0)]
    assert (2, 0) in oracle.dda(4, 0, 0, 0)


@pytest.fixture(scope="module")
def mutant_libs(oracle_lib, tmp_path_factory):
    d = tmp_path_factory.mktemp("mutants")
    libs = {}
    for m in MUTANTS:
        path = os.path.join(str(d), f"liboracle_m{m}.so")
        oracle.build_mutant(m, path)
        libs[m] = path
    return libs


@pytest.mark.parametrize("m", MUTANTS)
def test_each_planted_mistake_is_caught(mutant_libs, m):
    catching = [c for c in CASES if m in c["catches"]]
    assert catching, f"no case claims mutant {m}"
    with oracle.use_library(mutant_libs[m]):
        for c in catching:
            assert run_case(c) != [], f"mutant {m} passes {c['name']}"
