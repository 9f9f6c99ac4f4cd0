"""The oracle against the reference programs' own outputs (CPU only).

Fixtures were written by tests/golden/make_golden.py, which ran
score-matching and graph-alignment-optimized compiled unmodified from the
reference sources.
"""
import json
import os

import numpy as np
import pytest

import oracle_binding as ob

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _score_cases():
    with open(os.path.join(GOLD, "score_matching.json")) as f:
        return json.load(f)["cases"]


@pytest.mark.parametrize("case", _score_cases(), ids=lambda c: c["name"])
def test_score_ref_matches_score_matching(case):
    n = case["n"]
    o = ob.Oracle(n, case["g1"], case["g2"])
    mp = np.array(case["mapping"], np.int32)
    ref_map = np.where(mp > 0, mp, -1).astype(np.int32)  # unmapped -> not in the file
    ref_map[0] = -1
    assert o.score_ref(ref_map) == case["score"]


def _ga_dirs():
    return sorted(d for d in os.listdir(GOLD) if d.startswith("ga_opt_"))


def _load_edges(path):
    a = np.loadtxt(path, delimiter=",", skiprows=1, dtype=np.int64, ndmin=2)
    return tuple(a[:, i].astype(np.int32) for i in range(3))


def _load_mf(path, n):
    mp = np.zeros(n + 1, np.int32)
    with open(path) as f:
        next(f)
        for line in f:
            a, b = line.strip().split(",")
            mp[int(a[1:])] = int(b[1:])
    return mp


@pytest.mark.parametrize("name", _ga_dirs())
def test_first_improvement_matches_graph_alignment_optimized(name):
    d = os.path.join(GOLD, name)
    with open(os.path.join(d, "expected.json")) as f:
        exp = json.load(f)
    # vertex_range: the program's vertex range (NUM_NODES) where the run moved
    # images onto the empty vertices above the data
    n = exp.get("vertex_range", exp["n"])
    g1 = _load_edges(os.path.join(d, "male.csv"))
    g2 = _load_edges(os.path.join(d, "female.csv"))
    init = _load_mf(os.path.join(d, "init.csv"), n)
    o = ob.Oracle(n, g1, g2)
    if "initial_score" in exp:  # "Initial alignment score" (:720), lines leaving vertex 0 not scored
        assert o.score_opt(init) == exp["initial_score"]
    res = o.first_improvement(init, rows_per_sweep=exp["rows_per_sweep"], log_every=exp["log_every"])
    # the reference's progress samples, expanded from the stored change points
    changes = exp["log_changes"]
    last_row = min(exp["log_last_row"], int(res["log"][-1][0]))
    want, j = [], 0
    for row in range(exp["log_every"], last_row + 1, exp["log_every"]):
        while j + 1 < len(changes) and changes[j + 1][0] <= row:
            j += 1
        want.append((row, changes[j][1], changes[j][2]))
    got = [tuple(int(x) for x in r) for r in res["log"] if r[0] <= last_row]
    assert got == want
    np.testing.assert_array_equal(res["mapping"], np.array(exp["final_mapping"], np.int32))
    assert res["final_score"] == changes[-1][1]
