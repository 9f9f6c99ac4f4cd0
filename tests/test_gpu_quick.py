"""Changed-vertex evaluation, perturbation and the searches built on them, on
the GPU, against the oracle (quick.cu:141-250, 766-887, 1137-1324;
alignment-omp.c:495-541, 543-678).

bm_eval_rows and bm_eval_swaps_dense are compared entry by entry with the
oracle's Delta_B and Delta_G2 (the dense-matrix delta of quick.cu's kernels)
on graphs with every feature that separates the two; bm_perturb
draws the same vertices from the same generator state as the oracle and
leaves the device state consistent (score and rows after it); the two
drivers follow the oracle's runs step for step, so the final and best
mappings and every counter agree.
"""
import numpy as np
import pytest

import bmatch
import oracle_binding as ob
from conftest import random_mapping, random_multigraph

pytestmark = pytest.mark.gpu


def graphs(seed, n, m, dup=0.1, loops=0.05, zero_vertex=False):
    g1 = random_multigraph(seed, n, m, dup, loops, zero_vertex=zero_vertex)
    g2 = random_multigraph(seed + 100, n, m, dup, loops)
    return g1, g2


def case_graphs(case, n, m):
    if case == "negative":
        return random_multigraph(3, n, m, 0.1, 0.05, negative=True), random_multigraph(4, n, m, 0.1, 0.05)
    if case == "wide":  # weights past 2^15: the dense form reads them as short
        return random_multigraph(5, n, m, 0.1, 0.05, wmax=40000), random_multigraph(6, n, m, 0.1, 0.05, wmax=40000)
    return graphs(1, n, m, zero_vertex=case == "vertex0")


@pytest.mark.parametrize("kind", [bmatch.BM_DELTA_LISTS, bmatch.BM_DELTA_DENSE])
@pytest.mark.parametrize("case", ["sparse", "nonbijective", "vertex0", "negative", "wide"])
def test_eval_rows_matches_oracle(case, kind):
    n = 90
    g1, g2 = case_graphs(case, n, 900)
    mp = random_mapping(2, n, 0.1, bijective=case != "nonbijective")
    o = ob.Oracle(n, g1, g2)
    delta = o.delta_g2 if kind == bmatch.BM_DELTA_DENSE else o.delta_b
    rows = [1, 17, 45, n]
    moved = mp.copy()
    with bmatch.Engine(0) as e:
        e.upload(n, g1, g2)
        e.set_mapping(mp)
        got = e.eval_rows(rows, kind)
        # after moves: the lists stay, the diagonal follows the mapping
        for a, b in ((3, 8), (17, 60), (45, 2)):
            e.apply_swap(a, b)
            moved[a], moved[b] = moved[b], moved[a]
        got2 = e.eval_rows(rows, kind)
    for i, a in enumerate(rows):
        want = [0 if b == a or b == 0 else delta(mp, a, b) for b in range(n + 1)]
        np.testing.assert_array_equal(got[i], want, err_msg=f"{case} row {a}")
        want2 = [0 if b == a or b == 0 else delta(moved, a, b) for b in range(n + 1)]
        np.testing.assert_array_equal(got2[i], want2, err_msg=f"{case} row {a} after moves")


@pytest.mark.parametrize("case", ["multigraph", "vertex0", "negative", "wide"])
def test_eval_swaps_dense_matches_oracle(case):
    n = 40
    g1, g2 = case_graphs("sparse" if case == "multigraph" else case, n, 500)
    o = ob.Oracle(n, g1, g2)
    mp = random_mapping(7, n, 0.1, bijective=False)
    a, b = np.array([(x, y) for x in range(1, n + 1) for y in range(1, n + 1) if x != y], np.int32).T
    with bmatch.Engine(0) as e:
        # a first pair of graphs, then the real one: the dense-form lists follow the upload
        h1, h2 = graphs(90, n, 300)
        e.upload(n, h1, h2)
        e.set_mapping(mp)
        first = e.eval_swaps_dense(a[:50], b[:50])
        e.upload(n, g1, g2)
        e.set_mapping(mp)
        got = e.eval_swaps_dense(a, b)
        lists = e.eval_swaps(a, b)
        for bad in ((0, 1), (1, 1), (n + 1, 2)):
            with pytest.raises(bmatch.BMError):
                e.eval_swaps_dense([bad[0]], [bad[1]])
    o1 = ob.Oracle(n, h1, h2)
    assert first.tolist() == [o1.delta_g2(mp, int(x), int(y)) for x, y in zip(a[:50], b[:50])]
    want = np.array([o.delta_g2(mp, int(x), int(y)) for x, y in zip(a, b)], np.int32)
    np.testing.assert_array_equal(got, want)
    assert (got != lists).any()  # the case separates the two deltas


def test_eval_rows_errors():
    n = 20
    g1, g2 = graphs(5, n, 100)
    with bmatch.Engine(0) as e:
        e.upload(n, g1, g2)
        e.set_mapping(random_mapping(6, n))
        for bad in ([0], [n + 1], [3, 3], list(range(1, 66))):
            with pytest.raises(bmatch.BMError):
                e.eval_rows(bad)
        assert e.eval_rows([]).shape == (0, n + 1)
        for kind in (-1, 2):  # neither delta form
            with pytest.raises(bmatch.BMError):
                e.eval_rows([1], kind)
            with pytest.raises(bmatch.BMError):
                e.quick(max_evals=5, kind=kind)
            with pytest.raises(bmatch.BMError):
                e.global_best(max_steps=1, kind=kind)


@pytest.mark.parametrize("mode", [bmatch.BM_PERTURB_OMP, bmatch.BM_PERTURB_WEIGHTED])
def test_perturb_matches_oracle(mode):
    n = 70
    g1, g2 = graphs(7, n, 700)
    o = ob.Oracle(n, g1, g2)
    mp = random_mapping(8, n)
    st = 12345
    with bmatch.Engine(0) as e:
        e.upload(n, g1, g2)
        e.set_mapping(mp)
        cur = mp.copy()
        for k in (1, 3, 4, 10, 35):
            sel, st_dev = e.perturb(k, mode, st)
            cur, sel_o, st_o = o.perturb(cur, k, mode, st)
            assert sel.tolist() == sel_o.tolist() and st_dev == st_o, k
            st = st_dev
            np.testing.assert_array_equal(e.mapping(), cur)
            assert e.score() == o.score_opt(cur)
            # the device's patched state: rows after the perturbation
            rows = sorted({int(v) for v in sel if v >= 1})[:4]
            got = e.eval_rows(rows)
            for i, a in enumerate(rows):
                for b in (1, 2, n // 2, n):
                    if b != a:
                        assert got[i, b] == o.delta_b(cur, a, b)
        with pytest.raises(bmatch.BMError):
            e.perturb(36, mode, st)


def test_rng_matches_oracle():
    s1 = s2 = 42
    for _ in range(100):
        v1, s1 = bmatch.rng_next(s1)
        v2, s2 = ob.rng_next(s2)
        assert v1 == v2 and s1 == s2


@pytest.mark.parametrize("seed", [1, 2])
def test_row_best_restart_matches_oracle(seed):
    n = 60
    g1, g2 = graphs(20 + seed, n, 500, dup=0.0, loops=0.0)
    o = ob.Oracle(n, g1, g2)
    mp = random_mapping(30 + seed, n)
    kw = dict(seed=seed, initial_swaps=10, restart_swaps=8, idle_passes=2, max_passes=12)
    want_map, want = o.row_best_restart(mp, **kw)
    with bmatch.Engine(0) as e:
        e.upload(n, g1, g2)
        e.set_mapping(mp)
        best, st = e.row_best_restart(verify=True, **kw)
        np.testing.assert_array_equal(best, want_map)
        np.testing.assert_array_equal(e.mapping(), want_map)
        assert e.score() == want["best_score"]
    for f in ("passes", "restarts", "accepted", "initial_score", "final_score", "best_score"):
        assert st[f] == want[f], f
    assert st["restarts"] >= 1


@pytest.mark.parametrize("kind", [bmatch.BM_DELTA_DENSE, bmatch.BM_DELTA_LISTS])
@pytest.mark.parametrize("gap_norm", [5154247.0, 40.0])
def test_quick_matches_oracle(gap_norm, kind):
    n = 80
    g1, g2 = graphs(40, n, 800)
    o = ob.Oracle(n, g1, g2)
    mp = random_mapping(41, n)
    kw = dict(seed=9, max_evals=400, zero_limit=30, best_window=40, gap=0.002, gap_norm=gap_norm, kind=kind)
    want_map, want = o.quick(mp, **kw)
    with bmatch.Engine(0) as e:
        e.upload(n, g1, g2)
        e.set_mapping(mp)
        best, st = e.quick(**kw)
        np.testing.assert_array_equal(best, want_map)
        np.testing.assert_array_equal(e.mapping(), want_map)
        assert e.score() == want["best_score"]
    assert st == want
    assert st["positive_moves"] > 0 and st["zero_moves"] > 0 and st["perturbations"] > 1
