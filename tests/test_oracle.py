"""Pins of the CPU oracle that do not go through its own formulas (CPU only).

Each check uses an independent route: a plain numpy rescoring of the whole
alignment, closed-form corrections derived from the definitions, or a
textbook routine (itertools for permutation order).
"""
import itertools

import numpy as np
import pytest

import oracle_binding as ob
from conftest import random_mapping, random_multigraph


def np_score(n, g1, g2, mp):
    """S_opt by dense numpy: every G1 line with a source >= 1 (the score loop
    starts at vertex 1), W2 dense with the last line winning, as short."""
    W2 = np.zeros((n + 1, n + 1), np.int64)
    W2[g2[0], g2[1]] = g2[2]  # numpy fancy assignment keeps the last duplicate
    W2 = W2.astype(np.int16).astype(np.int64)
    keep = g1[0] >= 1
    return int(np.minimum(g1[2][keep], W2[mp[g1[0][keep]], mp[g1[1][keep]]]).sum())


def swapped(mp, a, b):
    m = mp.copy()
    m[a], m[b] = m[b], m[a]
    return m


def simple_graph(seed, n, m, w=30):
    """no self-loops, no repeated pairs"""
    rng = np.random.default_rng(seed)
    pairs = set()
    while len(pairs) < m:
        a, b = rng.integers(1, n + 1, 2)
        if a != b:
            pairs.add((int(a), int(b)))
    p = np.array(sorted(pairs), np.int32)
    return (p[:, 0], p[:, 1], rng.integers(1, w + 1, m).astype(np.int32))


def test_score_opt_is_numpy_rescore():
    for seed in range(6):
        zero = seed >= 4  # lines at vertex 0 in both graphs
        g1 = random_multigraph(seed, 30, 300, 0.2, 0.05, zero_vertex=zero)
        g2 = random_multigraph(seed + 50, 30, 300, 0.2, 0.05, zero_vertex=zero)
        mp = random_mapping(seed, 30, 0.1, bijective=seed % 2 == 0)
        if zero:
            mp[0] = 7  # a mapped vertex 0 would make its lines count if they were scored
            assert (g1[0] == 0).any()
        assert ob.Oracle(30, g1, g2).score_opt(mp) == np_score(30, g1, g2, mp)


def test_score_opt_skips_lines_leaving_vertex_0():
    """The three-line case: 0->1 (w 5) is not scored, 1->2 and 2->0 are."""
    g1 = (np.array([0, 1, 2], np.int32), np.array([1, 2, 0], np.int32), np.array([5, 4, 3], np.int32))
    g2 = (np.array([0, 1, 2], np.int32), np.array([1, 2, 0], np.int32), np.array([9, 9, 9], np.int32))
    o = ob.Oracle(2, g1, g2)
    assert o.score_opt(np.array([0, 1, 2], np.int32)) == 4 + 3


def test_score_invariants():
    g = simple_graph(1, 25, 120)
    o = ob.Oracle(25, g, g)
    ident = np.arange(26, dtype=np.int32)
    assert o.score_opt(ident) == int(g[2].sum())  # S(id; G, G) = sum of weights
    ref_ident = ident.copy()
    ref_ident[0] = -1
    assert o.score_ref(ref_ident) == int(g[2].sum())
    g2 = simple_graph(2, 25, 150)
    o2 = ob.Oracle(25, g, g2)
    for s in range(5):
        mp = random_mapping(s, 25)
        v = o2.score_opt(mp)
        assert 0 <= v <= min(int(g[2].sum()), int(g2[2].sum()))
        r = mp.copy()
        r[0] = -1
        assert o2.score_ref(r) == v  # no duplicates, positive weights: S_ref == S_opt


def test_delta_b_equals_rescore_without_loops_or_duplicates():
    n = 10
    g1, g2 = simple_graph(3, n, 45), simple_graph(4, n, 50)
    o = ob.Oracle(n, g1, g2)
    for s in range(3):
        mp = random_mapping(s, n, 0.2 if s == 2 else 0.0)
        for a in range(1, n + 1):
            for b in range(1, n + 1):
                if a != b:
                    want = np_score(n, g1, g2, swapped(mp, a, b)) - np_score(n, g1, g2, mp)
                    assert o.delta_b(mp, a, b) == want


def test_delta_b_self_loop_closed_form():
    """A self-loop (a,a,w) is scored by Delta_B as an out and an in entry of a
    (min(w,W2[fb][fa]) + min(w,W2[fa][fb]) - 2 min(w,W2[fa][fa])) instead of
    its true change min(w,W2[fb][fb]) - min(w,W2[fa][fa])."""
    n = 9
    base = simple_graph(5, n, 30)
    loops = np.array([2, 5, 7], np.int32)
    lw = np.array([4, 9, 2], np.int32)
    g1 = (np.concatenate([base[0], loops]), np.concatenate([base[1], loops]), np.concatenate([base[2], lw]))
    g2 = random_multigraph(6, n, 60, 0.0, 0.2)
    W2 = np.zeros((n + 1, n + 1), np.int64)
    W2[g2[0], g2[1]] = g2[2]
    o = ob.Oracle(n, g1, g2)
    mp = random_mapping(7, n)
    for a in range(1, n + 1):
        for b in range(1, n + 1):
            if a == b:
                continue
            exact = np_score(n, g1, g2, swapped(mp, a, b)) - np_score(n, g1, g2, mp)
            corr = 0
            for v, w in zip(loops, lw):
                if v in (a, b):
                    x, y = (a, b) if v == a else (b, a)
                    fx, fy = mp[x], mp[y]
                    corr += (min(w, W2[fy, fx]) + min(w, W2[fx, fy]) - 2 * min(w, W2[fx, fx])) - \
                            (min(w, W2[fy, fy]) - min(w, W2[fx, fx]))
            assert o.delta_b(mp, a, b) == exact + corr


def test_delta_g2_drops_self_loops_only():
    n = 9
    base = simple_graph(8, n, 30)
    g1 = (np.concatenate([base[0], [3]]), np.concatenate([base[1], [3]]), np.concatenate([base[2], [5]]))
    g2 = simple_graph(9, n, 40)
    o = ob.Oracle(n, g1, g2)
    ob_noloop = ob.Oracle(n, base, g2)
    mp = random_mapping(1, n)
    for a in range(1, n + 1):
        for b in range(a + 1, n + 1):
            # Delta_G2 never visits the loop, so it equals the loop-free graph's exact delta
            want = np_score(n, base, g2, swapped(mp, a, b)) - np_score(n, base, g2, mp)
            assert o.delta_g2(mp, a, b) == want == ob_noloop.delta_b(mp, a, b)


def dense_cell_score(n, g1, g2, mp):
    """Sum over the off-diagonal G1 cells (i, j >= 1) that are positive as short, last line winning, of
    min(cell, W2[p(i)][p(j)]): the quantity whose change quick.cu's dense kernel computes."""
    W1 = np.zeros((n + 1, n + 1), np.int64)
    W1[g1[0], g1[1]] = g1[2]
    W1 = W1.astype(np.int16).astype(np.int64)
    W2 = np.zeros((n + 1, n + 1), np.int64)
    W2[g2[0], g2[1]] = g2[2]
    W2 = W2.astype(np.int16).astype(np.int64)
    i, j = np.nonzero(W1 > 0)
    keep = (i != j) & (i >= 1) & (j >= 1)
    i, j = i[keep], j[keep]
    return int(np.minimum(W1[i, j], W2[mp[i], mp[j]]).sum())


@pytest.mark.parametrize("case", ["multigraph", "vertex0", "negative", "wide"])
def test_delta_g2_is_the_dense_cell_score_change(case):
    # repeated lines, self-loops, lines at vertex 0, weights <= 0 and weights that wrap as short: every
    # way the dense form departs from Delta_B, against an independent numpy rescore
    n = 12
    kw = dict(zero_vertex=case == "vertex0", negative=case == "negative", wmax=40000 if case == "wide" else 30)
    g1 = random_multigraph(21, n, 90, 0.3, 0.1, **kw)
    g2 = random_multigraph(22, n, 90, 0.3, 0.1, wmax=kw["wmax"])
    o = ob.Oracle(n, g1, g2)
    mp = random_mapping(23, n, 0.1, bijective=False)
    differs = 0
    for a in range(1, n + 1):
        for b in range(1, n + 1):
            if a == b:
                continue
            want = dense_cell_score(n, g1, g2, swapped(mp, a, b)) - dense_cell_score(n, g1, g2, mp)
            assert o.delta_g2(mp, a, b) == want, (a, b)
            differs += want != o.delta_b(mp, a, b)
    assert differs > 0  # the case does separate the two deltas


G3 = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]


def test_delta_g3_equals_rescore():
    n = 10
    g1, g2 = simple_graph(10, n, 50), simple_graph(11, n, 55)
    o = ob.Oracle(n, g1, g2)
    mp = random_mapping(2, n, 0.1)
    for t in itertools.combinations(range(1, n + 1), 3):
        for p in range(1, 6):
            m = mp.copy()
            f = [mp[t[0]], mp[t[1]], mp[t[2]]]
            for k in range(3):
                m[t[k]] = f[G3[p][k]]
            want = np_score(n, g1, g2, m) - np_score(n, g1, g2, mp)
            assert o.delta_g3(mp, *t, p) == want
            np.testing.assert_array_equal(ob.apply_g3(mp, *t, p), m)


def test_delta_p_is_exact_plus_intra_edges():
    """Delta_P counts every G1 line with both ends in the set twice."""
    n = 9
    g1 = random_multigraph(12, n, 70, 0.1, 0.1)
    g2 = random_multigraph(13, n, 70, 0.1, 0.1)
    W2 = np.zeros((n + 1, n + 1), np.int64)
    W2[g2[0], g2[1]] = g2[2]
    W2 = W2.astype(np.int16).astype(np.int64)
    o = ob.Oracle(n, g1, g2)
    mp = random_mapping(3, n)
    rng = np.random.default_rng(0)
    for _ in range(40):
        k = int(rng.integers(2, 6))
        s = rng.choice(np.arange(1, n + 1), k, replace=False).astype(np.int32)
        perm = rng.permutation(k).astype(np.int32)
        new = mp.copy()
        for i in range(k):
            new[s[i]] = mp[s[perm[i]]]
        exact = np_score(n, g1, g2, new) - np_score(n, g1, g2, mp)
        intra = 0
        for a, b, w in zip(*g1):
            if a in s and b in s:
                intra += min(w, W2[new[a], new[b]]) - min(w, W2[mp[a], mp[b]])
        assert o.delta_p(mp, s, perm) == exact + intra


def test_permutation_order_is_lexicographic():
    for k in range(1, 7):
        perms = list(itertools.permutations(range(k)))
        cur = np.arange(k, dtype=np.int32)
        for r, p in enumerate(perms):
            np.testing.assert_array_equal(cur, p)
            np.testing.assert_array_equal(ob.unrank_perm(k, r), p)
            ok, cur = ob.next_permutation(cur)
            assert ok == (r + 1 < len(perms))


def test_best_perm_strict_first_max():
    n = 8
    g1, g2 = random_multigraph(20, n, 60, 0.1, 0.05), random_multigraph(21, n, 60, 0.1, 0.05)
    o = ob.Oracle(n, g1, g2)
    mp = random_mapping(4, n)
    for s in ([1, 2, 3], [4, 6, 8, 2], [1, 3, 5, 7, 2]):
        s = np.array(s, np.int32)
        vals = [o.delta_p(mp, s, np.array(p, np.int32)) for p in itertools.permutations(range(len(s)))]
        d, r = o.best_perm(mp, s)
        best = max(vals)
        if best > 0:
            assert (d, r) == (best, vals.index(best))
        else:
            assert (d, r) == (0, -1)


def test_delta_a_is_out_edge_score_change():
    """Delta_A = change of the score restricted to out-lines of a and b that do
    not join a and b, over the de-duplicated graphs (last weight wins)."""
    n = 9
    g1 = random_multigraph(30, n, 60, 0.2, 0.1)
    g2 = random_multigraph(31, n, 60, 0.2, 0.0)
    d1 = {(int(a), int(b)): int(w) for a, b, w in zip(*g1)}
    d2 = {(int(a), int(b)): int(w) for a, b, w in zip(*g2)}
    o = ob.Oracle(n, g1, g2)
    mp = random_mapping(5, n, 0.2)

    def part(m, a, b):
        s = 0
        for (x, y), w in d1.items():
            if (x == a and y != b) or (x == b and y != a):
                if m[x] > 0 and m[y] > 0 and (m[x], m[y]) in d2:
                    s += min(w, d2[(m[x], m[y])])
        return s

    for a in range(1, n + 1):
        for b in range(1, n + 1):
            if a != b:
                assert o.delta_a(mp, a, b) == part(swapped(mp, a, b), a, b) - part(mp, a, b)


def test_first_improvement_brute_force():
    """The sweep applies a swap as soon as its delta is positive and resumes
    at the next m2 with the updated mapping."""
    n = 8
    g1, g2 = simple_graph(40, n, 25), simple_graph(41, n, 25)
    o = ob.Oracle(n, g1, g2)
    mp = random_mapping(6, n)
    cur, acc = mp.copy(), []
    while True:
        before = len(acc)
        for a in range(1, n + 1):
            for b in range(1, n + 1):
                if a == b:
                    continue
                d = np_score(n, g1, g2, swapped(cur, a, b)) - np_score(n, g1, g2, cur)
                if d > 0:
                    cur = swapped(cur, a, b)
                    acc.append((a, b, d))
        if len(acc) == before:
            break
    res = o.first_improvement(mp)
    np.testing.assert_array_equal(res["accepted"], np.array(acc, np.int32).reshape(-1, 3))
    np.testing.assert_array_equal(res["mapping"], cur)


def test_row_best_and_all_pairs_brute_force():
    n = 9
    g1, g2 = simple_graph(50, n, 30), simple_graph(51, n, 30)
    o = ob.Oracle(n, g1, g2)
    mp = random_mapping(8, n)
    D = {(a, b): np_score(n, g1, g2, swapped(mp, a, b)) - np_score(n, g1, g2, mp)
         for a in range(1, n + 1) for b in range(a + 1, n + 1)}
    r = o.all_pairs(mp, matrix=True)
    best = max(D.values())
    first = min(k for k, v in D.items() if v == best)
    assert r["best"] == ((*first, best) if best > 0 else (0, 0, 0))
    for (a, b), v in D.items():
        assert r["matrix"][a, b] == v == r["matrix"][b, a]
    # one row-best pass, brute force
    cur, acc = mp.copy(), []
    for a in range(1, n + 1):
        cands = [(np_score(n, g1, g2, swapped(cur, a, b)) - np_score(n, g1, g2, cur), -b)
                 for b in range(a + 1, n + 1)]
        if cands:
            d, nb = max(cands)
            if d > 0:
                cur = swapped(cur, a, -nb)
                acc.append((a, -nb, d))
    m_ref, acc_ref = o.row_best_pass(mp)
    np.testing.assert_array_equal(acc_ref, np.array(acc, np.int32).reshape(-1, 3))
    np.testing.assert_array_equal(m_ref, cur)


@pytest.mark.parametrize("threads", [1, 2])
def test_omp_rowbest_checksum_matches_sequential(threads):
    g1, g2 = simple_graph(60, 40, 300), simple_graph(61, 40, 300)
    o = ob.Oracle(40, g1, g2)
    mp = random_mapping(9, 40)
    ev, cs = o.omp_rowbest_rows(mp, 1, 41, threads)
    assert ev == 40 * 39 // 2
    r = o.all_pairs(mp)
    want = 0
    for a in range(1, 41):
        md, mv = -1, -1
        for b in range(a + 1, 41):
            d = o.delta_b(mp, a, b)
            if d > md:
                md, mv = d, b
        want += md
    assert cs == want
    assert r["best"][2] == max(0, max(r["row_best_delta"][1:40]))


def test_parallel_helpers_equal_the_single_routines():
    """The OpenMP helpers the full-size checks use against the routines they batch."""
    n = 40
    g1 = random_multigraph(21, n, 600, 0.1, 0.05, zero_vertex=True)
    g2 = random_multigraph(22, n, 600, 0.1, 0.05, zero_vertex=True)
    mp = random_mapping(23, n)
    o = ob.Oracle(n, g1, g2)
    ap = o.all_pairs(mp)
    rows = np.arange(1, n + 1, dtype=np.int32)
    b, d = o.row_best_rows(mp, rows, 4)
    np.testing.assert_array_equal(b, ap["row_best_b"][1:])
    np.testing.assert_array_equal(d, ap["row_best_delta"][1:])
    full = o.delta_rows(mp, np.array([3, 17, 40], np.int32), 4)
    for i, a in enumerate((3, 17, 40)):
        want = [0 if x in (0, a) else o.delta_b(mp, a, x) for x in range(n + 1)]
        assert full[i].tolist() == want
    # first-improvement rows: the sweep driver's first rows (a simple graph, where no rescore corrects a delta)
    s = simple_graph(24, n, 300)
    s2 = simple_graph(25, n, 320)
    os_ = ob.Oracle(n, s, s2)
    m0 = random_mapping(26, n)
    full_run = os_.first_improvement(m0, max_sweeps=1)
    rows_map, acc = os_.first_improvement_rows(m0, n, 4)
    np.testing.assert_array_equal(acc, full_run["accepted"])
    np.testing.assert_array_equal(rows_map, full_run["mapping"])
    rng = np.random.default_rng(27)
    tri = np.array([rng.choice(np.arange(1, n + 1), 3, replace=False) for _ in range(200)], np.int32)
    dd, pp = o.best_g3_many(mp, tri, 1, 4)
    assert [(int(x), int(y)) for x, y in zip(dd, pp)] == [o.best_g3(mp, *map(int, t), 1) for t in tri]
    sets = np.array([rng.choice(np.arange(1, n + 1), 5, replace=False) for _ in range(20)], np.int32)
    sd, sr = o.best_perm_many(mp, sets, 4)
    assert [(int(x), int(y)) for x, y in zip(sd, sr)] == [o.best_perm(mp, s_) for s_ in sets]
