"""Parity at the measured configurations, at their full sizes.

Each test builds the same seeded instance the benchmark measures and checks
the device against the oracle on it:

  C1  power-law pair n = 1,000 / 20,000 lines: first-improvement
      (graph-alignment-optimized.c:513-581) to convergence, every accepted
      swap in order, the final mapping and score.
  C2  Chung-Lu pair n = 18,524 / 3.0M lines, 30%-shuffled start: the
      all-pairs Delta_B sweep (:321-408) -- row bests of 200 leading and 100
      random rows, 50 whole rows, the global best (value, place, tie-break);
      the first 20 rows of the plain first-improvement sweep.
  C3  2^24 two-hop triples on the C1 pair: Delta_G3 (triple.cu:85-214) best
      rotation of a 50,000-triple sample, and the device arg-max over all.
  C4  4,096 8-vertex sets on the C1 pair: best permutation of 256 sets
      (graph-alignment-permset-parallel.c:304-339, next_permutation order),
      the device arg-max, and three batched permset steps replayed.
  C5  a served stream that includes an n = 20,000 job at 160 lines per
      vertex: its final score against S_opt, and the same result alone and
      in the mix.
"""
import numpy as np
import pytest

import bmatch
import oracle_binding as ob
from test_gpu_refine import replay, run as refine_run_logged

pytestmark = pytest.mark.gpu

THREADS = 16  # the oracle's OpenMP threads (the host share of one GPU)


@pytest.fixture(scope="module")
def c1():
    inst = bmatch.Instance.powerlaw(1, 1000, 20000)
    return inst, ob.Oracle(inst.n, inst.g1, inst.g2)


@pytest.fixture(scope="module")
def c2():
    inst = bmatch.Instance.chunglu(2, 18524, 3_000_000)
    print(f"[c2] n={inst.n} lines={len(inst.g1[0])}", flush=True)
    return inst, ob.Oracle(inst.n, inst.g1, inst.g2)


def test_c1_first_improvement_to_convergence(c1):
    inst, o = c1
    want = o.first_improvement(inst.p0)
    with bmatch.Engine(0) as e:
        e.upload(inst.n, inst.g1, inst.g2)
        e.set_mapping(inst.p0)
        assert e.score() == o.score_opt(inst.p0)
        acc, st = e.first_improvement(verify=True)
        np.testing.assert_array_equal(acc, want["accepted"])
        np.testing.assert_array_equal(e.mapping(), want["mapping"])
        assert st["final_score"] == want["final_score"] == e.score() == o.score_opt(want["mapping"])
        assert st["mismatches"] == 0 and st["sweeps"] == want["sweeps"]
        assert len(acc) > 1000
        # the last pass accepts nothing: once rows stop accepting (in that pass or at the end of the
        # one before) an all-pairs sweep finds no improving swap, and the rest is counted without
        # being searched
        assert 0 < st["rows_cleared"] < 2 * inst.n
        assert st["rows"] == st["sweeps"] * inst.n


def test_c2_all_pairs_sweep(c2):
    inst, o = c2
    n = inst.n
    rng = np.random.default_rng(5)
    with bmatch.Engine(0) as e:
        e.upload(n, inst.g1, inst.g2)
        e.set_mapping(inst.p0)
        got = e.sweep_all_pairs(rows=True)
        kind, _ = e.sweep_config()
        full_rows = np.sort(rng.choice(np.arange(1, n + 1), 50, replace=False)).astype(np.int32)
        gpu_rows = e.eval_rows(full_rows)
    assert kind == 2  # the sparse gain kernel the benchmark measures
    a, b, d = got["best"]
    rb_b, rb_d = got["row_best_b"], got["row_best_delta"]
    # the global best: its value is the oracle's delta, it tops every row, and it is the
    # lexicographically first pair with that value
    assert d > 0 and o.delta_b(inst.p0, a, b) == d
    assert d == rb_d[1:n].max()
    top = np.flatnonzero(rb_d[1:n] == d) + 1
    assert a == top[0] and b == rb_b[a]
    rows = np.concatenate([np.arange(1, 201), rng.choice(np.arange(201, n), 100, replace=False), [a]])
    rows = np.unique(rows).astype(np.int32)
    ob_b, ob_d = o.row_best_rows(inst.p0, rows, THREADS)
    np.testing.assert_array_equal(rb_d[rows], ob_d)
    np.testing.assert_array_equal(rb_b[rows], ob_b)
    assert rb_b[n] == 0 and rb_d[n] == np.iinfo(np.int32).min
    want_rows = o.delta_rows(inst.p0, full_rows, THREADS)
    np.testing.assert_array_equal(gpu_rows, want_rows)


def test_c2_first_improvement_rows(c2):
    inst, o = c2
    rows = 20
    mp_want, acc_want = o.first_improvement_rows(inst.p0, rows, THREADS)
    with bmatch.Engine(0) as e:
        e.upload(inst.n, inst.g1, inst.g2)
        e.set_mapping(inst.p0)
        s0 = e.score()
        acc, st = e.first_improvement(verify=True, max_rows=rows)
        np.testing.assert_array_equal(acc, acc_want)
        np.testing.assert_array_equal(e.mapping(), mp_want)
        assert st["mismatches"] == 0
        assert st["final_score"] == e.score() == o.score_opt(mp_want) == s0 + int(acc_want[:, 2].sum())
    assert len(acc_want) > 500


def test_c3_two_hop_triples(c1):
    inst, o = c1
    count = 1 << 24
    tri = inst.triples_2hop(7, count)
    sample = np.random.default_rng(3).choice(count, 50_000, replace=False)
    with bmatch.Engine(0) as e:
        e.upload(inst.n, inst.g1, inst.g2)
        e.set_mapping(inst.p0)
        d, p = e.eval_rot3(tri, bmatch.BM_ROT3_ROTATIONS)
        e.moves_set(3, tri)
        best = e.moves_eval(bmatch.BM_ROT3_ROTATIONS)
    want_d, want_p = o.best_g3_many(inst.p0, tri[sample], bmatch.BM_ROT3_ROTATIONS, THREADS)
    np.testing.assert_array_equal(d[sample], want_d)
    np.testing.assert_array_equal(p[sample], want_p)
    # the device arg-max over all 2^24: largest delta, lowest index on ties, from the per-triple results
    i = int(np.argmax(d))
    assert d[i] > 0 and best == (i, int(d[i]), int(p[i]))
    bd, bp = o.best_g3(inst.p0, *map(int, tri[i]), bmatch.BM_ROT3_ROTATIONS)
    assert (bd, bp) == (d[i], p[i])


def test_c4_permset_4096_sets(c1):
    inst, o = c1
    sets = bmatch.gen_sets(9, inst.n, 8, 4096)
    sample = np.sort(np.random.default_rng(4).choice(4096, 256, replace=False))
    with bmatch.Engine(0) as e:
        e.upload(inst.n, inst.g1, inst.g2)
        e.set_mapping(inst.p0)
        d, r = e.eval_permset(sets)
        e.moves_set(8, sets)
        best = e.moves_eval(0)
    want_d, want_r = o.best_perm_many(inst.p0, sets[sample], THREADS)
    np.testing.assert_array_equal(d[sample], want_d)
    np.testing.assert_array_equal(r[sample], want_r)
    i = int(np.argmax(d))
    assert d[i] > 0 and best == (i, int(d[i]), int(r[i]))


def test_c4_batched_permset_steps(c1):
    inst, o = c1
    prm = bmatch.refine_params(pairs=0, triples=0, sets=4096, set_k=8, set_gen=1, rounds=3, steps_per_check=1,
                               seed=1)
    with bmatch.Engine(0) as e:
        e.upload(inst.n, inst.g1, inst.g2)
        st, recs, hist = refine_run_logged(e, inst.p0, prm, max_steps=3, patience=4)
        assert st["steps"] == 3 and st["verified"] == 1 and st["set_moves"] > 0
        fin = replay(inst.n, inst.g1, o, inst.p0, recs, st, hist)
        np.testing.assert_array_equal(e.mapping(), fin)


def test_c5_served_stream_with_a_full_size_job():
    big = bmatch.Instance.powerlaw(777, 20000, 160 * 20000, shuffle_frac=0.3)
    small = [bmatch.Instance.powerlaw(780 + i, n, 160 * n, shuffle_frac=0.3) for i, n in enumerate((600, 3000, 9000))]
    prm = bmatch.refine_params(pairs=1, triples=1024, sets=0, rounds=3, steps_per_check=8, seed=11)
    kw = dict(move_budget=1_000_000, patience=8)
    with bmatch.Engine(0) as e:  # alone, through bm_refine_run
        e.upload(big.n, big.g1, big.g2)
        e.set_mapping(big.p0)
        e.refine_configure(prm)
        alone = e.refine_run(**kw)
        alone_map = e.mapping()
    assert alone["verified"] == 1 and alone["moves"] > 0
    with bmatch.Server(max_active=4, prep_threads=4) as srv:
        ids = [srv.submit(small[0].n, small[0].g1, small[0].g2, small[0].p0, prm, **kw)]
        ids.append(srv.submit(big.n, big.g1, big.g2, big.p0, prm, **kw))
        ids += [srv.submit(s.n, s.g1, s.g2, s.p0, prm, **kw) for s in small[1:]]
        res = [srv.wait(j, n) for j, n in zip(ids, (small[0].n, big.n, small[1].n, small[2].n))]
    for r, _ in res:
        assert r["status"] == 0, r["error"]
    r_big, m_big = res[1]
    assert r_big["n"] == big.n
    np.testing.assert_array_equal(m_big, alone_map)
    assert (r_big["final_score"], r_big["steps"], r_big["moves"]) == (alone["final_score"], alone["steps"],
                                                                       alone["moves"])
    o = ob.Oracle(big.n, big.g1, big.g2)
    assert o.score_opt(big.p0) == r_big["initial_score"]
    assert o.score_opt(m_big) == r_big["final_score"] > r_big["initial_score"]
