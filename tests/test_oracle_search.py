"""Pins of the oracle's perturbation, restart and quick-search restatements
(alignment-omp.c:495-541, 543-678; quick.cu:766-887, 1137-1324), CPU only.

The reference draws from a time-seeded rand(), so no run of it can be
replayed; these checks pin the generator against its published test vector,
the vertex weights against hand-computed cases, the draw rules against their
distributions, and the searches against invariants and a plain row-best
pass (which the restart loop reduces to when nothing triggers a restart).
"""
import numpy as np
import pytest

import oracle_binding as ob
from conftest import random_mapping, random_multigraph
from test_oracle import np_score, simple_graph


def test_splitmix64_published_vector():
    # Vigna's splitmix64.c seeded with 1234567 (the xoshiro/xorshift test vector)
    want = [6457827717110365317, 3203168211198807973, 9817491932198370423, 4593380528125082431,
            16408922859458223821]
    s = 1234567
    got = []
    for _ in range(5):
        v, s = ob.rng_next(s)
        got.append(v)
    assert got == want


def test_vertex_scores_hand_cases():
    # n = 4: lines 1->2 (5), 1->2 again (7: last wins), 2->1 (3), 3->3 loop (4),
    # 4->1 (-2: not positive), 0->1 (9: vertex 0 is outside 1..n), 2->3 (40000: wraps negative as short)
    g1 = (np.array([1, 1, 2, 3, 4, 0, 2]), np.array([2, 2, 1, 3, 1, 1, 3]), np.array([5, 7, 3, 4, -2, 9, 40000]))
    g2 = (np.array([1]), np.array([2]), np.array([1]))
    o = ob.Oracle(4, g1, g2)
    s = o.vertex_scores()
    f = np.float32
    want = [0.0,
            f(0.7) * f(7) + f(0.3) * f(3),  # v1: out {2:7}, in {2:3}
            f(0.7) * f(3) + f(0.3) * f(7),  # v2: out {1:3} (2->3 wraps), in {1:7}
            f(0.7) * f(4) + f(0.3) * f(4),  # v3: self-loop counts on both sides
            f(0.0)]                         # v4: only a negative line
    np.testing.assert_array_equal(s, np.array(want, np.float32))


def test_perturb_omp_draws_and_swaps():
    n = 40
    g1 = simple_graph(1, n, 200)
    o = ob.Oracle(n, g1, g1)
    m0 = random_mapping(2, n)
    st = 99
    for k in (1, 5, 20):
        m, sel, st2 = o.perturb(m0, k, 0, st)
        # the draws: the first 2k distinct values of next() % n
        want, s = [], st
        while len(want) < 2 * k:
            v, s = ob.rng_next(s)
            if v % n not in want:
                want.append(v % n)
        assert sel.tolist() == want and st2 == s
        exp = m0.copy()
        for i in range(k):
            exp[sel[i]], exp[sel[i + k]] = m0[sel[i + k]], m0[sel[i]]
        np.testing.assert_array_equal(m, exp)
        st = st2
    with pytest.raises(ValueError):
        o.perturb(m0, 21, 0, 1)


def test_perturb_weighted_follows_the_weights():
    # n = 10, weights from out-lines only: vertex v has one line v -> v%10+1 of weight v
    n = 10
    src = np.arange(1, n + 1)
    g1 = (src, src % n + 1, src.copy())
    o = ob.Oracle(n, g1, g1)
    sc = o.vertex_scores().astype(np.float64)
    p = sc[1:] / sc[1:].sum()
    m0 = np.arange(n + 1, dtype=np.int32)
    counts = np.zeros(n + 1)
    st = 7
    trials = 20000
    for _ in range(trials):
        _, sel, st = o.perturb(m0, 1, 1, st)
        counts[sel[0]] += 1
        assert sel[1] != sel[0] and 1 <= sel[1] <= n
    freq = counts[1:] / trials
    # binomial standard error at most 0.0035; 5 sigma
    assert np.all(np.abs(freq - p) < 0.018), (freq, p)


def test_perturb_weighted_without_replacement():
    n = 12
    g1 = simple_graph(3, n, 50)
    o = ob.Oracle(n, g1, g1)
    m0 = random_mapping(4, n)
    st = 5
    for _ in range(200):
        m, sel, st = o.perturb(m0, 6, 1, st)
        assert sorted(sel.tolist()) == list(range(1, n + 1))  # 2k = n: every vertex once
        assert sorted(m[1:].tolist()) == sorted(m0[1:].tolist())


def test_restart_without_restarts_is_row_best_passes():
    n = 50
    g1 = simple_graph(5, n, 400)
    g2 = simple_graph(6, n, 400)
    o = ob.Oracle(n, g1, g2)
    m0 = random_mapping(7, n)
    # idle_passes large: no restart; no initial perturbation; 2 passes
    got, st = o.row_best_restart(m0, seed=3, initial_swaps=0, restart_swaps=5, idle_passes=100, max_passes=2)
    ref = m0.copy()
    acc = np.zeros(3 * 10000, np.int32)
    a1 = ob.lib().orc_row_best_pass(o._h, ref, acc, 10000)
    s1 = o.score_opt(ref)
    best, best_map = s1, ref.copy()
    a2 = ob.lib().orc_row_best_pass(o._h, ref, acc, 10000)
    s2 = o.score_opt(ref)
    if s2 > best:
        best, best_map = s2, ref.copy()
    assert st["restarts"] == 0 and st["passes"] == 2 and st["accepted"] == a1 + a2
    assert st["best_score"] == best and st["final_score"] == s2
    # the returned mapping is the best one seen (the passes only raise the score here)
    assert o.score_opt(got) == best


def test_restart_keeps_the_best_and_restarts_when_idle():
    n = 30
    g1 = simple_graph(8, n, 150)
    g2 = simple_graph(9, n, 150)
    o = ob.Oracle(n, g1, g2)
    m0 = random_mapping(10, n)
    got, st = o.row_best_restart(m0, seed=11, initial_swaps=5, restart_swaps=4, idle_passes=2, max_passes=30)
    assert st["restarts"] >= 1
    assert st["initial_score"] == o.score_opt(m0)
    assert st["best_score"] == o.score_opt(got) == np_score(n, g1, g2, got)
    assert st["best_score"] >= st["initial_score"]
    assert sorted(got.tolist()) == sorted(m0.tolist())  # a permutation of the same images (vertex 0 included)


@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("gap_norm", [5154247.0, 50.0])
def test_quick_invariants(gap_norm, kind):
    n = 40
    g1 = random_multigraph(12, n, 300, 0.1, 0.05)
    g2 = random_multigraph(13, n, 300, 0.1, 0.05)
    o = ob.Oracle(n, g1, g2)
    m0 = random_mapping(14, n)
    got, st = o.quick(m0, seed=15, max_evals=300, zero_limit=20, best_window=30, gap_norm=gap_norm, kind=kind)
    assert st["evals"] == 300
    assert st["moves"] == st["positive_moves"] + st["zero_moves"]
    assert st["moves"] + st["perturbations"] == st["evals"]
    assert st["initial_score"] == o.score_opt(m0)
    assert st["best_score"] == o.score_opt(got) == np_score(n, g1, g2, got) >= st["initial_score"]
    assert sorted(got[1:].tolist()) == sorted(m0[1:].tolist()) and got[0] == m0[0]
    assert st["positive_moves"] > 0 and st["perturbations"] > 1


def test_quick_delta_kinds():
    # the two rankings coincide where the deltas do (no loops, no repeated lines, positive weights) and
    # part on a multigraph
    n = 40
    kw = dict(seed=5, max_evals=300, zero_limit=20, best_window=30)
    g1, g2 = simple_graph(30, n, 300), simple_graph(31, n, 300)
    o = ob.Oracle(n, g1, g2)
    m0 = random_mapping(32, n)
    a, sa = o.quick(m0, kind=0, **kw)
    b, sb = o.quick(m0, kind=1, **kw)
    assert sa == sb and (a == b).all()
    g1 = random_multigraph(33, n, 300, 0.3, 0.1)
    o = ob.Oracle(n, g1, g2)
    a, sa = o.quick(m0, kind=0, **kw)
    b, sb = o.quick(m0, kind=1, **kw)
    assert sa != sb
