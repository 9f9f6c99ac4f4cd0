"""The changed-vertex triple search of q3.cu on the CPU: the triple
enumeration and the oracle's restatement of the search loop.

The enumeration is pinned by a literal transcription of the kernel's index
arithmetic (q3.cu:59-150, an index -> triple map with while loops) against
the oracle's and the library's nested loops; the search by its invariants.
"""
import numpy as np
import pytest

import bmatch
import oracle_binding as ob
from conftest import random_mapping, random_multigraph


def kernel_unrank(n, chg):
    """q3.cu:59-150 for every thread index, with vertices_reorder = chg + the rest ascending."""
    k = len(chg)
    reorder = list(chg) + [v for v in range(1, n + 1) if v not in set(chg)]
    first = k * (k - 1) * (k - 2) // 6
    second = first + k * (k - 1) // 2 * (n - k)
    total = first + k * (k - 1) // 2 * (n - k) + k * (n - k - 1)
    out = []
    for idx in range(total):
        if idx < first:
            nn, a = k, 0
            t1 = (nn - 1) * (nn - 2) // 2
            t = idx
            while t >= t1:
                t -= t1
                a += 1
                nn -= 1
                t1 = (nn - 1) * (nn - 2) // 2
            b = a + 1
            t1 = nn - 2
            while t >= t1:
                t -= t1
                b += 1
                t1 -= 1
            c = b + t + 1
            out.append((reorder[a], reorder[b], reorder[c]))
        elif idx < second:
            i = idx - first
            pc = k * (k - 1) // 2
            c, rem = i // pc, i % pc
            a, t = 0, k - 1
            while rem >= t:
                rem -= t
                a += 1
                t -= 1
            b = a + rem + 1
            out.append((reorder[a], reorder[b], reorder[k + c]))
        else:
            i = idx - second
            rc = n - k - 1
            a, b = i // rc, i % rc
            out.append((reorder[a], reorder[k + b], reorder[k + b + 1]))
    return np.array(out, np.int32).reshape(-1, 3)


@pytest.mark.parametrize("n,chg", [(12, [3, 7, 9]), (20, [1, 2, 5, 11, 17, 20]), (30, [2, 4, 8, 9, 15, 16, 22, 30]),
                                   (9, [1, 2, 3, 4, 5, 6, 7, 8])])
def test_q3_enumeration_matches_the_kernel_index_map(n, chg):
    want = kernel_unrank(n, chg)
    np.testing.assert_array_equal(ob.q3_triples(n, chg), want)
    np.testing.assert_array_equal(bmatch.q3_triples(n, chg), want)
    assert len({tuple(sorted(t)) for t in want.tolist()}) == len(want)  # no triple twice


def test_q3_triples_rejects_unsorted_or_out_of_range():
    for chg in ([3, 2, 5], [0, 2, 5], [1, 2, 99]):
        with pytest.raises(bmatch.BMError):
            bmatch.q3_triples(20, chg)


def np_score(n, g1, g2, mp):
    W2 = np.zeros((n + 1, n + 1), np.int64)
    W2[g2[0], g2[1]] = g2[2]
    W2 = W2.astype(np.int16).astype(np.int64)
    keep = g1[0] >= 1
    return int(np.minimum(g1[2][keep], W2[mp[g1[0][keep]], mp[g1[1][keep]]]).sum())


@pytest.mark.parametrize("seed,dup,loops", [(1, 0.0, 0.0), (2, 0.15, 0.05)])
def test_q3_search_invariants(seed, dup, loops):
    n = 40
    g1 = random_multigraph(seed, n, 500, dup, loops, 30)
    g2 = random_multigraph(seed + 100, n, 500, dup, loops, 30)
    mp = random_mapping(seed + 7, n)
    o = ob.Oracle(n, g1, g2)
    best, st = o.q3(mp, seed=seed, max_evals=60)
    assert st["evals"] == 60 == st["perturbations"] + st["moves"]
    assert st["moves"] == st["positive_moves"] + st["zero_moves"]
    assert st["best_score"] == np_score(n, g1, g2, best) >= st["initial_score"] == np_score(n, g1, g2, mp)
    assert sorted(best[1:].tolist()) == sorted(mp[1:].tolist())
    assert st["triples"] > 0
    # the same seed gives the same run
    best2, st2 = o.q3(mp, seed=seed, max_evals=60)
    assert st2 == st and np.array_equal(best2, best)
