"""3-vertex (Delta_G3), k-set (Delta_P) and hill-climb (Delta_A) parity on the GPU."""
import numpy as np
import pytest

import bmatch
import oracle_binding as ob
from conftest import random_mapping, random_multigraph

pytestmark = pytest.mark.gpu


def setup(seed, n=40, m=400, dup=0.1, loops=0.05, unm=0.0):
    g1 = random_multigraph(seed, n, m, dup, loops, 40)
    g2 = random_multigraph(seed + 1, n, m, dup, loops, 40)
    mp = random_mapping(seed + 2, n, unm)
    e = bmatch.Engine(0)
    e.upload(n, g1, g2)
    e.set_mapping(mp)
    return n, g1, g2, mp, e, ob.Oracle(n, g1, g2)


@pytest.mark.parametrize("mode", [bmatch.BM_ROT3_ROTATIONS, bmatch.BM_ROT3_ALL])
@pytest.mark.parametrize("seed", [1, 2])
def test_rot3_matches_oracle(mode, seed):
    n, g1, g2, mp, e, o = setup(seed, unm=0.1)
    rng = np.random.default_rng(seed)
    tri = np.array([rng.choice(np.arange(1, n + 1), 3, replace=False) for _ in range(600)], np.int32)
    with e:
        d, p = e.eval_rot3(tri, mode)
    want = [o.best_g3(mp, *map(int, t), mode) for t in tri]
    np.testing.assert_array_equal(d, [w[0] for w in want])
    np.testing.assert_array_equal(p, [w[1] for w in want])


def test_rot3_each_permutation():
    # every p of triple.cu:85-92 through the ALL mode: the best over p in 1..5
    # must equal the max of the oracle's individual deltas
    n, g1, g2, mp, e, o = setup(9, n=20, m=300)
    rng = np.random.default_rng(9)
    tri = np.array([rng.choice(np.arange(1, n + 1), 3, replace=False) for _ in range(300)], np.int32)
    with e:
        d, p = e.eval_rot3(tri, bmatch.BM_ROT3_ALL)
    for t, dd, pp in zip(tri, d, p):
        vals = [o.delta_g3(mp, *map(int, t), q) for q in range(1, 6)]
        best = max([0] + vals)
        assert dd == best
        if best > 0:
            assert pp == 1 + vals.index(best)


def test_apply_rot3():
    n, g1, g2, mp, e, o = setup(4)
    cur = mp.copy()
    with e:
        for p in range(1, 6):
            e.apply_rot3(3, 7, 11, p)
            cur = ob.apply_g3(cur, 3, 7, 11, p)
            np.testing.assert_array_equal(e.mapping(), cur)
            assert e.score() == o.score_opt(cur)


@pytest.mark.parametrize("k", [2, 3, 4, 5, 8])
def test_permset_matches_oracle(k):
    n, g1, g2, mp, e, o = setup(10 + k, n=30, m=350)
    sets = bmatch.gen_sets(100 + k, n, k, 12 if k == 8 else 60)
    with e:
        d, r = e.eval_permset(sets)
    for s, dd, rr in zip(sets, d, r):
        wd, wr = o.best_perm(mp, s)
        assert (dd, rr) == (wd, wr), (s, dd, rr, wd, wr)


def test_permset_rank_decoding_and_apply():
    n, g1, g2, mp, e, o = setup(31, n=30, m=350)
    sets = bmatch.gen_sets(7, n, 4, 40)
    cur = mp.copy()
    with e:
        d, r = e.eval_permset(sets)
        for s, dd, rr in zip(sets, d, r):
            if rr < 0:
                continue
            perm = bmatch.perm_from_rank(4, int(rr))
            np.testing.assert_array_equal(perm, ob.unrank_perm(4, int(rr)))
            assert o.delta_p(cur, s, perm) == dd
            e.apply_perm(s, perm)
            cur = ob.apply_perm(cur, s, perm)
            np.testing.assert_array_equal(e.mapping(), cur)
            assert e.score() == o.score_opt(cur)
            break


def test_hillclimb_delta_matches_oracle():
    # Delta_A: de-duplicated G1 out-lists; G2 weights in [1, 32767] so the
    # dense short cell and the hash-table lookup agree
    n, g1, g2, mp, e, o = setup(41, n=35, m=400, unm=0.15)
    a, b = np.nonzero(~np.eye(n + 1, dtype=bool))
    with e:
        got = e.eval_swaps_hillclimb(a, b)
    want = [o.delta_a(mp, int(x), int(y)) for x, y in zip(a, b)]
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("k", [2, 3, 4, 8])
def test_resident_moves_match_host_lists(k):
    n, g1, g2, mp, e, o = setup(50 + k, n=40, m=450, unm=0.1)
    rng = np.random.default_rng(k)
    cnt = 30 if k == 8 else 700
    moves = np.array([rng.choice(np.arange(1, n + 1), k, replace=False) for _ in range(cnt)], np.int32)
    with e:
        e.moves_set(k, moves)
        for mode in ([bmatch.BM_ROT3_ROTATIONS, bmatch.BM_ROT3_ALL] if k == 3 else [0]):
            bi, bd, bx = e.moves_eval(mode)
            d, x = e.moves_results(cnt)
            if k == 2:
                want_d = e.eval_swaps(moves[:, 0], moves[:, 1])
                want_x = np.zeros(cnt, np.int32)
            elif k == 3:
                want_d, want_x = e.eval_rot3(moves, mode)
            else:
                want_d, want_x = e.eval_permset(moves)
            np.testing.assert_array_equal(d, want_d)
            if k >= 3:
                np.testing.assert_array_equal(x, want_x)
            if (want_d > 0).any():
                i = int(np.argmax(want_d))  # first index of the maximum
                assert (bi, bd, bx) == (i, int(want_d[i]), int(want_x[i]))
            else:
                assert (bi, bd, bx) == (-1, 0, 0)
        # the list stays resident across a move
        a, b = int(moves[0][0]), int(moves[0][1])
        e.apply_swap(a, b)
        mp2 = mp.copy()
        mp2[a], mp2[b] = mp2[b], mp2[a]
        e.moves_eval(bmatch.BM_ROT3_ALL if k == 3 else 0)
        d2, _ = e.moves_results(cnt)
        if k == 2:
            np.testing.assert_array_equal(d2, [o.delta_b(mp2, int(p), int(q)) for p, q in moves])


@pytest.mark.parametrize("lanes", [8, 16, 32, 64])
def test_rot3_every_lane_width(lanes, monkeypatch):
    # 4,096 triples or more go through the autotuner, which the variable overrides
    monkeypatch.setenv("BM_ROT3_LANES", str(lanes))
    n, g1, g2, mp, e, o = setup(31, unm=0.05)
    rng = np.random.default_rng(lanes)
    tri = np.array([rng.choice(np.arange(1, n + 1), 3, replace=False) for _ in range(4200)], np.int32)
    with e:
        d, p = e.eval_rot3(tri, bmatch.BM_ROT3_ALL)
        assert e.refine_config()[0] == lanes
    want = [o.best_g3(mp, *map(int, t), bmatch.BM_ROT3_ALL) for t in tri]
    np.testing.assert_array_equal(d, [w[0] for w in want])
    np.testing.assert_array_equal(p, [w[1] for w in want])


@pytest.mark.parametrize("case", ["multigraph", "vertex0", "negative", "g2negative", "wide", "unmapped"])
@pytest.mark.parametrize("table", [0, 1])
def test_rot3_table_form_matches_oracle(case, table, monkeypatch):
    # the table form (all-pairs pass over the dense-form lists + 3x3 blocks per triple) and the direct kernel,
    # each forced, on every G1 feature the dense scan of triple.cu treats specially; both call paths (host
    # list, resident list), both modes, and again after a move changed the mapping
    monkeypatch.setenv("BM_ROT3_TABLE", str(table))
    n = 45
    kw = dict(zero_vertex=case == "vertex0", negative=case == "negative", wmax=40000 if case == "wide" else 40)
    g1 = random_multigraph(70, n, 600, 0.15, 0.05, **kw)
    # g2negative: small positive G1 weights (the packed partial sums of the all-pairs pass would apply)
    # against negative G2 cells
    g2 = random_multigraph(71, n, 600, 0.15, 0.05, wmax=kw["wmax"], negative=case == "g2negative")
    mp = random_mapping(72, n, 0.15 if case == "unmapped" else 0.0, bijective=case != "unmapped")
    o = ob.Oracle(n, g1, g2)
    rng = np.random.default_rng(73)
    tri = np.array([rng.choice(np.arange(1, n + 1), 3, replace=False) for _ in range(5000)], np.int32)
    with bmatch.Engine(0) as e:
        e.upload(n, g1, g2)
        e.set_mapping(mp)
        got = {mode: e.eval_rot3(tri, mode) for mode in (bmatch.BM_ROT3_ROTATIONS, bmatch.BM_ROT3_ALL)}
        e.moves_set(3, tri[:700])
        e.moves_eval(bmatch.BM_ROT3_ALL)
        res = e.moves_results(700)
        e.apply_swap(int(tri[0][0]), int(tri[0][1]))
        mp2 = e.mapping()
        e.moves_eval(bmatch.BM_ROT3_ALL)
        res2 = e.moves_results(700)
    for mode, (d, p) in got.items():
        want = [o.best_g3(mp, *map(int, t), mode) for t in tri]
        np.testing.assert_array_equal(d, [w[0] for w in want], err_msg=f"mode {mode}")
        np.testing.assert_array_equal(p, [w[1] for w in want], err_msg=f"mode {mode}")
    want = [o.best_g3(mp, *map(int, t), bmatch.BM_ROT3_ALL) for t in tri[:700]]
    np.testing.assert_array_equal(res[0], [w[0] for w in want])
    np.testing.assert_array_equal(res[1], [w[1] for w in want])
    want = [o.best_g3(mp2, *map(int, t), bmatch.BM_ROT3_ALL) for t in tri[:700]]
    np.testing.assert_array_equal(res2[0], [w[0] for w in want])
    np.testing.assert_array_equal(res2[1], [w[1] for w in want])
    assert (got[bmatch.BM_ROT3_ALL][0] > 0).any()
