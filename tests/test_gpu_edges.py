"""Degenerate inputs on the GPU against the oracle: graphs without lines,
the smallest n, a star with isolated vertices, an all-unmapped matching and
a matching that sends many vertices to one image."""
import numpy as np
import pytest

import bmatch
import oracle_binding as ob
from conftest import random_mapping, random_multigraph

pytestmark = pytest.mark.gpu

NONE = (np.zeros(0, np.int32),) * 3


def check_all(n, g1, g2, mp, refine=True):
    o = ob.Oracle(n, g1, g2)
    with bmatch.Engine(0) as e:
        e.upload(n, g1, g2)
        e.set_mapping(mp)
        assert e.score() == o.score_opt(mp)
        if n >= 2:
            a = np.array([1 + i % n for i in range(3 * n)], np.int32)
            b = np.array([1 + (i * 7 + 1) % n for i in range(3 * n)], np.int32)
            b = np.where(a == b, b % n + 1, b)
            want = [o.delta_b(mp, int(x), int(y)) for x, y in zip(a, b)]
            np.testing.assert_array_equal(e.eval_swaps(a, b), want)
        sw = e.sweep_all_pairs(rows=True, matrix=True)
        ref = o.all_pairs(mp, matrix=True)
        assert sw["best"] == ref["best"]
        np.testing.assert_array_equal(sw["matrix"], ref["matrix"])
        rows = list(range(1, min(n, 4) + 1))
        R = e.eval_rows(rows)
        for i, r in enumerate(rows):
            want = [0 if c in (0, r) else o.delta_b(mp, r, c) for c in range(n + 1)]
            np.testing.assert_array_equal(R[i], want)
        for r in rows:
            for mode in (bmatch.BM_ROW_FIRST_POSITIVE, bmatch.BM_ROW_ARGMAX):
                got = e.eval_swap_row(r, 1, n, mode)
                ds = [(c, o.delta_b(mp, r, c)) for c in range(1, n + 1) if c != r]
                if mode == bmatch.BM_ROW_FIRST_POSITIVE:
                    pos = [x for x in ds if x[1] > 0]
                    assert got == (pos[0] if pos else (0, 0))
                elif ds:
                    best = max(d for _, d in ds)
                    assert got == (min(c for c, d in ds if d == best), best)
        acc, st = e.first_improvement(verify=True)
        want = o.first_improvement(mp)
        np.testing.assert_array_equal(acc, want["accepted"])
        assert st["final_score"] == want["final_score"]
        if refine and n >= 3:
            e.set_mapping(mp)
            e.refine_configure(bmatch.refine_params(pairs=1, triples=64, sets=16, set_k=3, rounds=2,
                                                    steps_per_check=2, seed=1))
            rs = e.refine_run(max_steps=6, patience=2)
            assert rs["verified"] == 1
            assert rs["final_score"] == e.score() == o.score_opt(e.mapping())


@pytest.mark.parametrize("which", ["g1", "g2", "both"])
def test_graphs_without_lines(which):
    n = 30
    g = random_multigraph(1, n, 200)
    g1 = NONE if which in ("g1", "both") else g
    g2 = NONE if which in ("g2", "both") else random_multigraph(2, n, 200)
    check_all(n, g1, g2, random_mapping(3, n))


def test_negative_lines_against_an_empty_g2():
    n = 25
    g1 = random_multigraph(4, n, 150, negative=True)
    check_all(n, g1, NONE, random_mapping(5, n))


@pytest.mark.parametrize("n", [1, 2, 3, 4])
def test_smallest_n(n):
    g1 = (np.array([1, n], np.int32), np.array([n, 1], np.int32), np.array([5, 3], np.int32))
    g2 = (np.array([1, n], np.int32), np.array([n, 1], np.int32), np.array([4, 9], np.int32))
    mp = np.arange(n + 1, dtype=np.int32)[::-1].copy()
    mp[0] = 0
    mp[1:] = np.arange(n, 0, -1)
    check_all(n, g1, g2, mp, refine=n >= 3)


def test_refine_rejects_too_small_n():
    g1 = (np.array([1], np.int32), np.array([2], np.int32), np.array([1], np.int32))
    with bmatch.Engine(0) as e:
        e.upload(2, g1, g1)
        e.set_mapping(np.array([0, 2, 1], np.int32))
        with pytest.raises(bmatch.BMError):
            e.refine_configure(bmatch.refine_params(pairs=1))


def test_star_and_isolated_vertices():
    n = 40
    hub = 7
    leaves = np.array([v for v in range(1, 31) if v != hub], np.int32)
    g1 = (np.concatenate([np.full(len(leaves), hub), leaves[:10]]).astype(np.int32),
          np.concatenate([leaves, np.full(10, hub)]).astype(np.int32),
          np.arange(1, len(leaves) + 11, dtype=np.int32))
    g2 = random_multigraph(6, n, 120)
    check_all(n, g1, g2, random_mapping(7, n))


def test_all_unmapped_and_collisions():
    n = 35
    g1 = random_multigraph(8, n, 300, 0.1, 0.05)
    g2 = random_multigraph(9, n, 300, 0.1, 0.05)
    check_all(n, g1, g2, np.zeros(n + 1, np.int32), refine=False)
    many = np.zeros(n + 1, np.int32)
    many[1:] = 5  # every vertex on one image
    many[::4] = 9
    many[0] = 0
    check_all(n, g1, g2, many, refine=False)
