"""The batched lexicographic triple pass of triple.cu on the CPU: the triple
order, and the oracle's restatement of the pass loop.

The order is pinned by itertools.combinations (the lexicographic order the
reference's batch index intends, triple.cu:54-75) for the oracle and the
library; the pass by a brute-force restatement in Python on small simple
graphs, where Delta_G3 is the exact rescore difference, with the splitmix64
draws of the library generator (itself pinned by its published vector).
"""
import itertools

import numpy as np
import pytest

import bmatch
import oracle_binding as ob
from conftest import random_mapping, random_multigraph

# triple.cu:85-92: the images (f1, f2, f3) after permutation p
PERMS = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]


@pytest.mark.parametrize("n", [3, 5, 9, 14])
def test_lex_triples_is_combinations_order(n):
    want = np.array(list(itertools.combinations(range(1, n + 1), 3)), np.int32).reshape(-1, 3)
    total = len(want)
    for start, count in [(0, total), (0, 1), (total - 1, 5), (total // 3, 7), (total // 2, total)]:
        end = min(total, start + count)
        np.testing.assert_array_equal(ob.lex_triples(n, start, count), want[start:end])
        np.testing.assert_array_equal(bmatch.lex_triples(n, start, count), want[start:end])
    assert len(bmatch.lex_triples(n, total, 4)) == 0


def test_lex_triples_rejects_bad_arguments():
    for n, start, count in [(2, 0, 1), (10, -1, 1), (10, 0, -1)]:
        with pytest.raises(bmatch.BMError):
            bmatch.lex_triples(n, start, count)


def np_score(n, g1, g2, mp):
    W2 = np.zeros((n + 1, n + 1), np.int64)
    W2[g2[0], g2[1]] = g2[2]
    W2 = W2.astype(np.int16).astype(np.int64)
    keep = g1[0] >= 1
    return int(np.minimum(g1[2][keep], W2[mp[g1[0][keep]], mp[g1[1][keep]]]).sum())


def brute_pass(n, g1, g2, mp, seed, max_evals, batch):
    """triple.cu:941-1196 by rescoring every permutation of every triple."""
    triples = list(itertools.combinations(range(1, n + 1), 3))
    total = len(triples)
    cur_map = np.array(mp, np.int32)
    best_map = cur_map.copy()
    cur = best = init = np_score(n, g1, g2, cur_map)
    rng = seed
    evals = passes = perts = moves = count = 0
    while evals < max_evals:
        found, start = False, 0
        passes += 1
        while start < total and evals < max_evals:
            chunk = triples[start:start + batch]
            evals += 1
            count += len(chunk)
            bd, bt, bp = 0, None, 0
            for t in chunk:
                f = [int(cur_map[v]) for v in t]
                d_best, p_best = 0, 0
                for p in range(1, 6):
                    m = cur_map.copy()
                    for i in range(3):
                        m[t[i]] = f[PERMS[p][i]]
                    d = np_score(n, g1, g2, m) - cur
                    if d > d_best:
                        d_best, p_best = d, p
                if d_best > bd:
                    bd, bt, bp = d_best, t, p_best
            if bd > 0:
                f = [int(cur_map[v]) for v in bt]
                for i in range(3):
                    cur_map[bt[i]] = f[PERMS[bp][i]]
                cur += bd
                moves += 1
                if cur > best:
                    best, best_map = cur, cur_map.copy()
                found = True
                break
            start += batch
        if not found and start >= total:
            cur_map = best_map.copy()
            v, rng = bmatch.rng_next(rng)
            k = 3 + v % 3
            sel = []
            while len(sel) < 2 * k:
                v, rng = bmatch.rng_next(rng)
                if v % n not in sel:
                    sel.append(v % n)
            orig = [int(cur_map[s]) for s in sel]
            for i in range(k):
                cur_map[sel[i]] = orig[i + k]
                cur_map[sel[i + k]] = orig[i]
            cur = np_score(n, g1, g2, cur_map)
            perts += 1
    return best_map, {"evals": evals, "passes": passes, "perturbations": perts, "moves": moves,
                      "initial_score": init, "final_score": cur, "best_score": best, "triples": count}


def simple_graph(seed, n, m, wmax=20):
    """m distinct lines s -> d, s != d, in 1..n (no repeated pair, self-loop or vertex 0)."""
    rng = np.random.default_rng(seed)
    pairs = [(s, d) for s in range(1, n + 1) for d in range(1, n + 1) if s != d]
    pick = rng.choice(len(pairs), m, replace=False)
    src = np.array([pairs[i][0] for i in pick], np.int32)
    dst = np.array([pairs[i][1] for i in pick], np.int32)
    return src, dst, rng.integers(1, wmax + 1, m).astype(np.int32)


@pytest.mark.parametrize("seed", [1, 2])
def test_triples_pass_matches_brute_force(seed):
    n = 11
    g1 = simple_graph(seed, n, 50)
    g2 = simple_graph(seed + 50, n, 50)
    mp = random_mapping(seed + 9, n)
    o = ob.Oracle(n, g1, g2)
    got_map, got = o.triples_pass(mp, seed=seed, max_evals=40, batch=40)
    want_map, want = brute_pass(n, g1, g2, mp, seed, 40, 40)
    assert got == want
    np.testing.assert_array_equal(got_map, want_map)
    assert got["moves"] > 0 and got["perturbations"] > 0


@pytest.mark.parametrize("seed,simple", [(3, True), (4, False)])
def test_triples_pass_invariants(seed, simple):
    n = 30
    if simple:
        g1, g2 = simple_graph(seed, n, 300, 30), simple_graph(seed + 100, n, 300, 30)
    else:
        g1 = random_multigraph(seed, n, 400, 0.15, 0.05, 30)
        g2 = random_multigraph(seed + 100, n, 400, 0.15, 0.05, 30)
    mp = random_mapping(seed + 7, n)
    o = ob.Oracle(n, g1, g2)
    best, st = o.triples_pass(mp, seed=seed, max_evals=50, batch=900)
    assert st["evals"] == 50 and st["initial_score"] == np_score(n, g1, g2, mp)
    if simple:  # Delta_G3 is the exact change, so the running score is the score
        assert st["best_score"] == np_score(n, g1, g2, best) >= st["initial_score"]
    # the images are permuted among vertices 0..n (a perturbation can draw index 0)
    assert sorted(best.tolist()) == sorted(mp.tolist())
    best2, st2 = o.triples_pass(mp, seed=seed, max_evals=50, batch=900)
    assert st2 == st and np.array_equal(best2, best)
