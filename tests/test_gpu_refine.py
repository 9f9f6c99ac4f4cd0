"""Batched refinement on the GPU: every applied step is checked on the CPU.

For each step of a run the log of applied moves is replayed on the oracle,
round by round: the moves of one round share no vertex and no G1 line, no
move shares a vertex with a move of an earlier round of the step, each
move's logged delta equals the oracle's exact score change when it is
applied to the mapping the earlier rounds left (its logged base delta is the
change at the step's starting mapping: for swaps Delta_B of
graph-alignment-optimized.c:321-408 and the arg-max of its row over b > a),
and the step's moves together raise S_opt by exactly their sum, which is the
score the device recorded for that step.
"""
import numpy as np
import pytest

import bmatch
import oracle_binding as ob
from conftest import random_mapping, random_multigraph

pytestmark = pytest.mark.gpu


def closed_nbrs(n, g1):
    nb = [{v} for v in range(n + 1)]
    for s, d in zip(g1[0].tolist(), g1[1].tolist()):
        nb[s].add(d)
        nb[d].add(s)
    return nb


def simple_vertices(n, g1):
    """Vertices no self-loop, repeated line or line from vertex 0 touches."""
    ok = np.ones(n + 1, bool)
    ok[0] = False
    seen = set()
    for s, d in zip(g1[0].tolist(), g1[1].tolist()):
        if s == d or s == 0:
            ok[d] = False
        if (s, d) in seen:
            ok[s] = ok[d] = False
        seen.add((s, d))
    return ok


def sorted_log(recs):
    """Log records (step, kind, k, delta, round, base, vertices[8], images[8]) by step, round, first vertex."""
    order = np.lexsort((recs[:, 6], recs[:, 4], recs[:, 0]))
    return recs[order]


def replay(n, g1, o, mp0, recs, stats, hist, check_rows=True):
    """Check the whole run against the oracle; returns the final mapping."""
    nb = closed_nbrs(n, g1)
    simple = simple_vertices(n, g1)
    cur = mp0.copy()
    score = o.score_opt(cur)
    assert score == stats["initial_score"]
    kinds = [0, 0, 0]
    for step in range(stats["steps"]):
        ms = recs[recs[:, 0] == step]
        taken = set()
        nxt, now, gain = cur.copy(), score, 0
        for rnd in sorted(set(ms[:, 4].tolist())):
            before, before_score = nxt.copy(), now
            in_round = set()
            for r in ms[ms[:, 4] == rnd]:
                kind, k, delta, base = int(r[1]), int(r[2]), int(r[3]), int(r[5])
                vs = [int(x) for x in r[6:6 + k]]
                img = [int(x) for x in r[14:14 + k]]
                kinds[kind] += 1
                assert delta > 0
                assert all(1 <= v <= n and simple[v] for v in vs)
                assert not (set(vs) & taken), f"step {step}: move {vs} reuses a vertex of an earlier round"
                assert sorted(img) == sorted(int(cur[v]) for v in vs)
                closed = set().union(*(nb[v] for v in vs))
                assert not (closed & in_round), f"step {step} round {rnd}: move {vs} conflicts with an earlier one"
                in_round.update(vs)
                # the exact change on top of the earlier rounds, and the candidate's delta at the step's start
                moved = before.copy()
                moved[vs] = img
                assert o.score_opt(moved) - before_score == delta, (step, rnd, kind, vs)
                alone = cur.copy()
                alone[vs] = img
                assert o.score_opt(alone) - score == base, (step, rnd, kind, vs)
                if not taken:
                    assert delta == base
                if kind == 0:
                    a, b = vs
                    assert a < b and o.delta_b(cur, a, b) == base
                    if check_rows:
                        row = [o.delta_b(cur, a, x) for x in range(a + 1, n + 1)]
                        assert max(row) == base and row.index(base) == b - a - 1
                nxt[vs] = img
                gain += delta
            taken |= in_round
            now = o.score_opt(nxt)
            assert now == before_score + sum(int(r[3]) for r in ms[ms[:, 4] == rnd])
        h = hist[step]
        assert h[:4] == (now, gain, len(ms), 1), (step, h)
        assert h[4] in (0, sum(int(r[2]) for r in ms))  # moved vertices (counted with the incremental update)
        cur, score = nxt, now
    assert score == stats["final_score"]
    assert kinds == [stats["pair_moves"], stats["triple_moves"], stats["set_moves"]]
    return cur


def run(e, mp0, params, log_cap=1 << 16, **kw):
    e.set_mapping(mp0)
    e.refine_configure(params)
    e.refine_log(log_cap)
    st = e.refine_run(**kw)
    recs, cnt = e.refine_moves()
    assert cnt == len(recs) == st["moves"]
    hist = e.refine_history(0, st["steps"])
    return st, sorted_log(recs), hist


def instance(seed, n, m):
    inst = bmatch.Instance.powerlaw(seed, n, m)
    return inst, ob.Oracle(inst.n, inst.g1, inst.g2)


def test_pairs_only_converges_exactly():
    inst, o = instance(11, 150, 1500)
    with bmatch.Engine(0) as e:
        e.upload(inst.n, inst.g1, inst.g2)
        st, recs, hist = run(e, inst.p0, bmatch.refine_params(pairs=1, rounds=2, steps_per_check=1), patience=1)
        assert st["verified"] == 1 and st["converged"] == 1 and st["moves"] > 0
        fin = replay(inst.n, inst.g1, o, inst.p0, recs, st, hist)
        np.testing.assert_array_equal(e.mapping(), fin)
        assert e.score() == st["final_score"] == o.score_opt(fin)
        # converged: no improving swap is left (the generator makes every vertex simple)
        assert e.sweep_all_pairs()["best"][2] <= 0


def test_later_rounds_score_on_top_of_earlier_winners():
    # power-law pair: the first rounds' winners hold hubs whose neighbourhoods cover most vertices, so
    # later rounds take moves next to them, each scored given what the earlier rounds moved
    inst, o = instance(41, 300, 6000)
    prm = bmatch.refine_params(pairs=1, triples=1000, sets=100, set_k=4, rounds=12, steps_per_check=2, seed=3)
    with bmatch.Engine(0) as e:
        e.upload(inst.n, inst.g1, inst.g2)
        st, recs, hist = run(e, inst.p0, prm, max_steps=6, patience=6)
        assert st["verified"] == 1 and st["moves"] > 0
        fin = replay(inst.n, inst.g1, o, inst.p0, recs, st, hist)
        np.testing.assert_array_equal(e.mapping(), fin)
        nb = closed_nbrs(inst.n, inst.g1)
        # some move of a later round touches a vertex an earlier round of its step moved
        adjacent = 0
        for step in range(st["steps"]):
            ms = recs[recs[:, 0] == step]
            for r in ms:
                earlier = ms[ms[:, 4] < r[4]]
                near = set().union(*(nb[int(v)] for v in r[6:6 + int(r[2])]))
                adjacent += any(int(v) in near for q in earlier for v in q[6:6 + int(q[2])])
        assert adjacent > 0
        assert (recs[:, 3] != recs[:, 5]).any()


@pytest.mark.parametrize("k", [3, 4, 8])
def test_all_kinds_exact(k):
    inst, o = instance(20 + k, 200, 2200)
    prm = bmatch.refine_params(pairs=1, triples=3000, sets=60 if k == 8 else 400, set_k=k, rounds=3,
                               steps_per_check=2, seed=5 + k)
    with bmatch.Engine(0) as e:
        e.upload(inst.n, inst.g1, inst.g2)
        st, recs, hist = run(e, inst.p0, prm, max_steps=12, patience=4)
        assert st["verified"] == 1 and st["moves"] > 0
        assert st["triple_moves"] > 0 and st["set_moves"] > 0
        fin = replay(inst.n, inst.g1, o, inst.p0, recs, st, hist, check_rows=False)
        np.testing.assert_array_equal(e.mapping(), fin)


def test_nonsimple_vertices_never_move():
    n = 60
    g1 = random_multigraph(3, n, 700, dup_frac=0.05, selfloop_frac=0.05)
    g2 = random_multigraph(4, n, 700, dup_frac=0.05, selfloop_frac=0.05)
    mp = random_mapping(5, n, 0.1)
    o = ob.Oracle(n, g1, g2)
    with bmatch.Engine(0) as e:
        e.upload(n, g1, g2)
        prm = bmatch.refine_params(pairs=1, triples=2000, sets=200, set_k=4, rounds=2, steps_per_check=3, seed=9)
        st, recs, hist = run(e, mp, prm, max_steps=9, patience=3)
        assert st["verified"] == 1
        fin = replay(n, g1, o, mp, recs, st, hist, check_rows=False)
        np.testing.assert_array_equal(e.mapping(), fin)


def test_deterministic_and_variant_independent(monkeypatch):
    inst, o = instance(31, 180, 2000)
    prm = bmatch.refine_params(pairs=1, triples=2000, sets=300, set_k=5, rounds=2, steps_per_check=2, seed=77)
    results = []
    # triples from the lists (direct, each lane width) or from the step's X (the default on this pair,
    # whose lines are all their own dense cells), with either sweep kernel
    for lanes, threads, kernel, rot3 in [(None, None, None, None), (8, 64, "dense", "direct"),
                                         (32, 256, "sparse", "direct"), (16, 64, "sparse", "direct"),
                                         (None, 64, "dense", None), (None, 256, "sparse", None)]:
        for name, val in (("BM_ROT3_LANES", lanes), ("BM_PERMSET_THREADS", threads), ("BM_SWEEP_KERNEL", kernel),
                          ("BM_REFINE_ROT3", rot3)):
            if val is None:
                monkeypatch.delenv(name, raising=False)
            else:
                monkeypatch.setenv(name, str(val))
        with bmatch.Engine(0) as e:
            e.upload(inst.n, inst.g1, inst.g2)
            st, recs, hist = run(e, inst.p0, prm, max_steps=8, patience=8)
            if lanes:
                assert e.refine_config() == (lanes, threads)
            if kernel:
                assert e.sweep_config()[0] == (1 if kernel == "dense" else 2)
            results.append((st["final_score"], st["moves"], recs.tobytes(), e.mapping().tobytes()))
    assert all(r == results[0] for r in results[1:])


@pytest.mark.parametrize("k", [4, 8])
def test_incremental_x_update_matches_full_sweeps(k, monkeypatch):
    # X kept current between steps (moved rows, columns and corrections only)
    # must give the same steps as recomputing it whole every step
    inst, o = instance(60 + k, 220, 2600)
    prm = bmatch.refine_params(pairs=1, triples=1500, sets=200, set_k=k, rounds=2, steps_per_check=2, seed=19 + k)
    runs = []
    # X and P / P^T both updated in place (rows of X gathered per row in LDS, or per comb entry with
    # global atomics); X in place with P rebuilt; both rebuilt
    for inc_max, patch_max, xdelta in (("100000", "100000", "rows"), ("100000", "100000", "entries"),
                                       ("100000", "0", "rows"), ("0", "0", "rows")):
        monkeypatch.setenv("BM_REFINE_INC_MAX", inc_max)
        monkeypatch.setenv("BM_REFINE_PATCH_MAX", patch_max)
        monkeypatch.setenv("BM_REFINE_XDELTA", xdelta)
        with bmatch.Engine(0) as e:
            e.upload(inst.n, inst.g1, inst.g2)
            st, recs, hist = run(e, inst.p0, prm, max_steps=14, patience=6)
            assert st["verified"] == 1 and st["moves"] > 0
            fin = replay(inst.n, inst.g1, o, inst.p0, recs, st, hist, check_rows=(inc_max != "0"))
            np.testing.assert_array_equal(e.mapping(), fin)
            runs.append((recs.tobytes(), e.mapping().tobytes(), st["final_score"]))
    assert all(r == runs[0] for r in runs[1:])


def test_incremental_update_with_loops_dups_and_unmapped(monkeypatch):
    monkeypatch.setenv("BM_REFINE_INC_MAX", "100000")
    monkeypatch.setenv("BM_REFINE_PATCH_MAX", "100000")
    n = 70
    g1 = random_multigraph(13, n, 900, dup_frac=0.05, selfloop_frac=0.05)
    g2 = random_multigraph(14, n, 900, dup_frac=0.05, selfloop_frac=0.05)
    mp = random_mapping(15, n, 0.1)
    o = ob.Oracle(n, g1, g2)
    with bmatch.Engine(0) as e:
        e.upload(n, g1, g2)
        prm = bmatch.refine_params(pairs=1, triples=800, sets=100, set_k=4, rounds=2, steps_per_check=3, seed=4)
        st, recs, hist = run(e, mp, prm, max_steps=12, patience=6)
        assert st["verified"] == 1
        fin = replay(n, g1, o, mp, recs, st, hist, check_rows=False)
        np.testing.assert_array_equal(e.mapping(), fin)


def test_packed_runs_equal_solo_runs():
    # graph pairs of uneven n in one launch per chunk: each ends exactly as alone
    specs = [(71, 90, 900, 2), (72, 260, 3000, 3), (73, 170, 1800, 2)]
    prms = [bmatch.refine_params(pairs=1, triples=700, sets=80 * (i % 2), set_k=4, rounds=2, steps_per_check=spc,
                                 seed=5 + i) for i, (_, _, _, spc) in enumerate(specs)]
    insts = [bmatch.Instance.powerlaw(sd, n, m) for sd, n, m, _ in specs]
    solo = []
    for inst, prm in zip(insts, prms):
        with bmatch.Engine(0) as e:
            e.upload(inst.n, inst.g1, inst.g2)
            e.set_mapping(inst.p0)
            e.refine_configure(prm)
            st = e.refine_run(max_steps=40, patience=4)
            solo.append((st, e.mapping()))
    engines = []
    try:
        for inst, prm in zip(insts, prms):
            e = bmatch.Engine(0)
            engines.append(e)
            e.upload(inst.n, inst.g1, inst.g2)
            e.set_mapping(inst.p0)
            e.refine_configure(prm)
        packed = bmatch.refine_run_many(engines, max_steps=40, patience=4)
        for inst, e, st, (sst, smap) in zip(insts, engines, packed, solo):
            assert st == sst
            np.testing.assert_array_equal(e.mapping(), smap)
            assert e.score() == st["final_score"] == ob.Oracle(inst.n, inst.g1, inst.g2).score_opt(smap)
    finally:
        for e in engines:
            e.close()


def test_runs_continue_from_the_current_mapping():
    inst, o = instance(41, 160, 1700)
    with bmatch.Engine(0) as e:
        e.upload(inst.n, inst.g1, inst.g2)
        e.set_mapping(inst.p0)
        e.refine_configure(bmatch.refine_params(pairs=1, triples=1000, rounds=2, steps_per_check=1, seed=3))
        a = e.refine_run(max_steps=3, patience=5)
        mid = e.mapping()
        assert a["final_score"] == o.score_opt(mid)
        b = e.refine_run(max_steps=3, patience=5)
        assert a["steps"] == b["steps"] == 3
        assert b["initial_score"] == a["final_score"]
        assert e.score() == b["final_score"] == o.score_opt(e.mapping())
        # a host move between runs is picked up by the next run
        e.apply_swap(1, 2)
        moved = e.mapping()
        c = e.refine_run(max_steps=1, patience=5)
        assert c["initial_score"] == o.score_opt(moved)
        assert e.score() == c["final_score"] == o.score_opt(e.mapping())
        # and so is a new mapping
        e.set_mapping(inst.p0)
        d = e.refine_run(max_steps=1, patience=5)
        assert d["initial_score"] == o.score_opt(inst.p0)


def test_back_to_back_runs_equal_one_run():
    # a run that starts from the mapping the previous one left keeps its gain matrix
    inst, o = instance(81, 240, 2800)
    prm = bmatch.refine_params(pairs=1, triples=600, sets=60, set_k=4, rounds=2, steps_per_check=3, seed=8)
    ends = []
    for parts in ((6,), (3, 3), (2, 2, 2)):
        with bmatch.Engine(0) as e:
            e.upload(inst.n, inst.g1, inst.g2)
            e.set_mapping(inst.p0)
            e.refine_configure(prm)
            e.refine_log(1 << 14)
            for k in parts:
                st = e.refine_run(max_steps=k, patience=10)
                assert st["verified"] == 1 and st["steps"] == k
                assert st["final_score"] == o.score_opt(e.mapping())
            hist = e.refine_history(0, parts[-1])
            if len(parts) > 1:
                assert all(h[5] == 0 for h in hist)  # no whole sweep: X carried over
            ends.append((e.mapping().tobytes(), st["final_score"]))
    assert ends[0] == ends[1] == ends[2]


def test_move_budget_stops_at_check():
    inst, o = instance(51, 200, 2400)
    with bmatch.Engine(0) as e:
        e.upload(inst.n, inst.g1, inst.g2)
        e.set_mapping(inst.p0)
        e.refine_configure(bmatch.refine_params(pairs=1, rounds=2, steps_per_check=1))
        st = e.refine_run(move_budget=5, patience=3)
        assert st["moves"] >= 5
        hist = e.refine_history(0, st["steps"])
        assert sum(h[2] for h in hist[:-1]) < 5  # stopped at the first check past the budget
        assert st["final_score"] == o.score_opt(e.mapping())


def test_refine_errors():
    inst, o = instance(61, 50, 400)
    with bmatch.Engine(0) as e:
        e.upload(inst.n, inst.g1, inst.g2)
        with pytest.raises(bmatch.BMError):
            e.refine_configure(bmatch.refine_params())  # no mapping yet
        e.set_mapping(inst.p0)
        with pytest.raises(bmatch.BMError):
            e.refine_run()  # not configured
        with pytest.raises(bmatch.BMError):
            e.refine_configure(bmatch.refine_params(pairs=0))
        with pytest.raises(bmatch.BMError):
            e.refine_configure(bmatch.refine_params(sets=10, set_k=9))
        with pytest.raises(bmatch.BMError):
            e.refine_configure(bmatch.refine_params(rounds=-1))
        with pytest.raises(bmatch.BMError):
            e.refine_configure(bmatch.refine_params(rounds=65))
        e.refine_configure(bmatch.refine_params(rounds=0))  # 0: chosen by size
