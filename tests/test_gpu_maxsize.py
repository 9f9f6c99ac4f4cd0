"""The largest graphs the engine takes: n = 65,534 (the translated G2 lists
keep vertex ids in 16 bits, 0xFFFF marks an empty slot).  W2, P and the gain
matrix X are 17.2 GB each here, so every device index past 2^31 elements is
exercised; the sparse sweep's LDS row no longer fits, so the dense gain
kernels run.  Checked against the oracle: the score, pair deltas, whole rows,
the sweep's best pair and row bests, a move, and batched steps.
"""
import numpy as np
import pytest

import bmatch
import oracle_binding as ob
from conftest import random_mapping, random_multigraph

pytestmark = pytest.mark.gpu

N_MAX = 65534


@pytest.fixture(scope="module")
def big():
    n = N_MAX
    g1 = random_multigraph(61, n, 200_000, 0.05, 0.01)
    g2 = random_multigraph(62, n, 200_000, 0.05, 0.01)
    mp = random_mapping(63, n, 0.05)
    return n, g1, g2, mp, ob.Oracle(n, g1, g2)


def test_max_n_deltas_rows_and_sweep(big):
    n, g1, g2, mp, o = big
    rng = np.random.default_rng(64)
    with bmatch.Engine(0) as e:
        e.upload(n, g1, g2)
        e.set_mapping(mp)
        assert e.score() == o.score_opt(mp)
        a = rng.integers(1, n + 1, 3000)
        b = rng.integers(1, n + 1, 3000)
        b = np.where(a == b, b % n + 1, b)
        got = e.eval_swaps(a, b)
        want = [o.delta_b(mp, int(x), int(y)) for x, y in zip(a, b)]
        np.testing.assert_array_equal(got, want)
        # whole rows, including the last vertex and rows far apart in the tile grid
        rows = [1, 4097, n // 2, n]
        R = e.eval_rows(rows)
        for i, r in enumerate(rows):
            cols = np.concatenate([rng.integers(1, n + 1, 2000), [1, n, r - 1 if r > 1 else 2]])
            for c in cols:
                c = int(c)
                assert R[i, c] == (0 if c == r else o.delta_b(mp, r, c)), (r, c)
        sw = e.sweep_all_pairs(rows=True)
        ba, bb, bd = sw["best"]
        rb, rd = sw["row_best_b"], sw["row_best_delta"]
        if bd > 0:
            assert o.delta_b(mp, ba, bb) == bd
        assert bd == max(0, int(rd[1:n].max()))
        # row bests agree with the whole rows (arg-max over b > a, lowest b on ties)
        for i, r in enumerate(rows):
            if r == n:
                continue
            seg = R[i, r + 1:]
            assert rd[r] == seg.max() and rb[r] == r + 1 + int(np.argmax(seg))
        # a move, then the score and a row again
        e.apply_swap(rows[0], rows[2])
        m2 = mp.copy()
        m2[rows[0]], m2[rows[2]] = m2[rows[2]], m2[rows[0]]
        assert e.score() == o.score_opt(m2)
        R2 = e.eval_rows([rows[2]])
        for c in rng.integers(1, n + 1, 500):
            c = int(c)
            assert R2[0, c] == (0 if c == rows[2] else o.delta_b(m2, rows[2], c))


def test_max_n_batched_steps(big):
    n, g1, g2, mp, o = big
    with bmatch.Engine(0) as e:
        e.upload(n, g1, g2)
        e.set_mapping(mp)
        e.refine_configure(bmatch.refine_params(pairs=1, triples=2048, sets=64, set_k=4, rounds=2,
                                                steps_per_check=2, seed=3))
        st = e.refine_run(max_steps=2, patience=2)
        assert st["verified"] == 1 and st["moves"] > 0
        assert st["initial_score"] == o.score_opt(mp)
        assert st["final_score"] == e.score() == o.score_opt(e.mapping())


@pytest.mark.parametrize("n,kind", [(36855, 2), (36856, 1)])
def test_sparse_row_limit(n, kind):
    # the sparse sweep keeps one X row in LDS (4 (n+1) bytes <= 144 KiB - 32): the last n that
    # fits runs it, the next one falls back to the dense slab kernels; both agree with the oracle
    g1 = random_multigraph(71, n, 150_000)
    g2 = random_multigraph(72, n, 150_000)
    mp = random_mapping(73, n)
    o = ob.Oracle(n, g1, g2)
    rng = np.random.default_rng(74)
    with bmatch.Engine(0) as e:
        e.upload(n, g1, g2)
        e.set_mapping(mp)
        sw = e.sweep_all_pairs(rows=True)
        assert e.sweep_config()[0] == kind
        ba, bb, bd = sw["best"]
        rb, rd = sw["row_best_b"], sw["row_best_delta"]
        assert bd > 0 and o.delta_b(mp, ba, bb) == bd and bd == int(rd[1:n].max())
        rows = [int(r) for r in rng.integers(1, n, 4)] + [ba]
        R = e.eval_rows(sorted(set(rows)))
        for i, r in enumerate(sorted(set(rows))):
            seg = R[i, r + 1:]
            assert rd[r] == seg.max() and rb[r] == r + 1 + int(np.argmax(seg))
            for c in rng.integers(1, n + 1, 300):
                c = int(c)
                assert R[i, c] == (0 if c == r else o.delta_b(mp, r, c)), (r, c)
