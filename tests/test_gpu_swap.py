"""Delta_B parity on the GPU: every swap kernel against the oracle.

Bar: bit-exact integers (graph-alignment-optimized.c:321-408), including
self-loops, repeated CSV lines (dense cell = last line), unmapped vertices
(mapping 0), non-bijective mappings and vertex 0.
"""
import numpy as np
import pytest

import bmatch
import oracle_binding as ob
from conftest import random_mapping, random_multigraph

pytestmark = pytest.mark.gpu

CASES = [
    # (name, n, m, dup, loops, unmapped, bijective, zero_vertex)
    ("simple", 40, 300, 0.0, 0.0, 0.0, True, False),
    ("dups_loops", 50, 500, 0.2, 0.05, 0.0, True, False),
    ("unmapped", 45, 400, 0.1, 0.05, 0.2, True, False),
    ("nonbijective", 30, 250, 0.1, 0.05, 0.1, False, False),
    ("vertex0", 35, 300, 0.1, 0.05, 0.0, True, True),
    ("dense", 24, 400, 0.3, 0.1, 0.0, True, False),
    ("negative", 30, 300, 0.1, 0.05, 0.1, True, False),  # weights < 0: unpacked int32 path
    ("wide", 40, 300, 0.1, 0.05, 0.0, True, False),      # weights up to 30000: unpacked int32 path
]


def make_case(case, seed=7):
    name, n, m, dup, loops, unm, bij, zero = case
    wmax = 30000 if name == "wide" else 30
    g1 = random_multigraph(seed, n, m, dup, loops, wmax, zero, negative=name == "negative")
    g2 = random_multigraph(seed + 1, n, m, dup, loops, wmax, zero, negative=name == "negative")
    mp = random_mapping(seed + 2, n, unm, bij)
    if zero:
        mp[0] = 3  # the reference lets a mapping line name vertex 0
    return n, g1, g2, mp


def engine_for(n, g1, g2, mp):
    e = bmatch.Engine(0)
    e.upload(n, g1, g2)
    e.set_mapping(mp)
    return e


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_score_and_pairs(case):
    n, g1, g2, mp = make_case(case)
    o = ob.Oracle(n, g1, g2)
    with engine_for(n, g1, g2, mp) as e:
        assert e.score() == o.score_opt(mp)
        a, b = np.nonzero(~np.eye(n + 1, dtype=bool))
        got = e.eval_swaps(a, b)
        want = np.array([o.delta_b(mp, x, y) for x, y in zip(a, b)], np.int32)
        np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("case", CASES[:4], ids=[c[0] for c in CASES[:4]])
def test_p_rebuild_one_row_per_block(case, monkeypatch):
    # the P / P^T rebuild for rows too wide to stage twice in LDS
    monkeypatch.setenv("BM_GATHER_SPLIT", "1")
    n, g1, g2, mp = make_case(case, seed=41)
    o = ob.Oracle(n, g1, g2)
    with engine_for(n, g1, g2, mp) as e:
        assert e.score() == o.score_opt(mp)
        ref = o.all_pairs(mp, matrix=True)
        got = e.sweep_all_pairs(matrix=True)
        np.testing.assert_array_equal(got["matrix"], ref["matrix"])


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_all_pairs_matrix(case):
    n, g1, g2, mp = make_case(case)
    o = ob.Oracle(n, g1, g2)
    ref = o.all_pairs(mp, matrix=True)
    with engine_for(n, g1, g2, mp) as e:
        got = e.sweep_all_pairs(rows=True, matrix=True)
    np.testing.assert_array_equal(got["matrix"], ref["matrix"])
    assert got["best"] == ref["best"]
    np.testing.assert_array_equal(got["row_best_b"][1:n], ref["row_best_b"][1:n])
    np.testing.assert_array_equal(got["row_best_delta"][1:n], ref["row_best_delta"][1:n])


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_row_kernel(case):
    n, g1, g2, mp = make_case(case)
    o = ob.Oracle(n, g1, g2)
    rng = np.random.default_rng(3)
    with engine_for(n, g1, g2, mp) as e:
        for a in range(0, n + 1):
            lo = int(rng.integers(0, n + 1))
            hi = int(rng.integers(lo, n + 1))
            ds = {b: o.delta_b(mp, a, b) for b in range(lo, hi + 1) if b != a}
            pos = [b for b in sorted(ds) if ds[b] > 0]
            want_fp = (pos[0], ds[pos[0]]) if pos else (0, 0)
            assert e.eval_swap_row(a, lo, hi, bmatch.BM_ROW_FIRST_POSITIVE) == want_fp, (a, lo, hi)
            if ds:
                bb = max(ds, key=lambda b: (ds[b], -b))
                want_am = (bb, ds[bb])
            else:
                want_am = (0, np.iinfo(np.int32).min)
            assert e.eval_swap_row(a, lo, hi, bmatch.BM_ROW_ARGMAX) == want_am, (a, lo, hi)


def test_apply_swap_patches_state():
    n, g1, g2, mp = make_case(CASES[1])
    o = ob.Oracle(n, g1, g2)
    rng = np.random.default_rng(11)
    cur = mp.copy()
    with engine_for(n, g1, g2, mp) as e:
        for _ in range(25):
            a, b = rng.choice(np.arange(1, n + 1), 2, replace=False)
            e.apply_swap(int(a), int(b))
            cur[a], cur[b] = cur[b], cur[a]
            np.testing.assert_array_equal(e.mapping(), cur)
            assert e.score() == o.score_opt(cur)
            x, y = rng.integers(1, n + 1, 64), rng.integers(1, n + 1, 64)
            keep = x != y
            got = e.eval_swaps(x[keep], y[keep])
            want = [o.delta_b(cur, int(p), int(q)) for p, q in zip(x[keep], y[keep])]
            np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_row_group_kernel(case):
    """bm_eval_swap_rows: rows a..a+count-1 under one mapping, the first over [lo, n], the
    others over [lo_next, n] ([r+1, n] for lo_next < 0), each as the oracle's row search."""
    n, g1, g2, mp = make_case(case)
    o = ob.Oracle(n, g1, g2)
    rng = np.random.default_rng(5)
    imin = np.iinfo(np.int32).min
    with engine_for(n, g1, g2, mp) as e:
        groups = [(1, 16), (n - 15, 16), (n - 2, 3), (n, 1), (0, 5)]
        groups += [(int(rng.integers(0, n - 6)), int(rng.integers(1, 8))) for _ in range(6)]
        for a, count in groups:
            lo = int(rng.integers(0, n + 2))
            for mode in (bmatch.BM_ROW_FIRST_POSITIVE, bmatch.BM_ROW_ARGMAX):
                for lo_next in (1, -1, 0):
                    bs, ds_ = e.eval_swap_rows(a, count, lo, lo_next, mode)
                    for i in range(count):
                        r = a + i
                        rlo = lo if i == 0 else (r + 1 if lo_next < 0 else lo_next)
                        ds = {b: o.delta_b(mp, r, b) for b in range(max(rlo, 0), n + 1) if b != r}
                        if mode == bmatch.BM_ROW_FIRST_POSITIVE:
                            pos = [b for b in sorted(ds) if ds[b] > 0]
                            want = (pos[0], ds[pos[0]]) if pos else (0, 0)
                        else:
                            bb = max(ds, key=lambda b: (ds[b], -b)) if ds else 0
                            want = (bb, ds[bb]) if ds else (0, imin)
                        assert (int(bs[i]), int(ds_[i])) == want, (a, count, lo, lo_next, mode, i)
        for bad in ((1, 0), (1, bmatch.BM_ROW_GROUP_MAX + 1), (n - 2, 4), (-1, 2)):
            with pytest.raises(bmatch.BMError):
                e.eval_swap_rows(bad[0], bad[1], 1, 1)


def test_first_improvement_trace_matches_oracle():
    inst = bmatch.Instance.powerlaw(5, 120, 1500)
    o = ob.Oracle(inst.n, inst.g1, inst.g2)
    ref = o.first_improvement(inst.p0)
    with engine_for(inst.n, inst.g1, inst.g2, inst.p0) as e:
        acc, st = e.first_improvement(verify=True)
        assert st["mismatches"] == 0
        np.testing.assert_array_equal(acc, ref["accepted"])
        np.testing.assert_array_equal(e.mapping(), ref["mapping"])
        assert st["final_score"] == ref["final_score"] == o.score_opt(ref["mapping"])
        assert st["sweeps"] == ref["sweeps"]


def test_first_improvement_with_loops_and_dups():
    n, g1, g2, mp = make_case(CASES[1], seed=21)
    o = ob.Oracle(n, g1, g2)
    ref = o.first_improvement(mp, max_sweeps=3)
    with engine_for(n, g1, g2, mp) as e:
        acc, st = e.first_improvement(max_sweeps=3, verify=True)
        np.testing.assert_array_equal(acc, ref["accepted"])
        np.testing.assert_array_equal(e.mapping(), ref["mapping"])
        assert st["final_score"] == ref["final_score"]


@pytest.mark.parametrize("unmapped,bijective", [(0.2, True), (0.1, False)], ids=["unmapped", "nonbijective"])
def test_first_improvement_dense_route_to_convergence(unmapped, bijective):
    """Mappings the sparse kernels do not take (unmapped or colliding images): rows and the
    quiet-stretch sweep go through the dense kernels; the moves stay the oracle's."""
    n = 160
    g1 = random_multigraph(31, n, 1600, 0.1, 0.05, 30, False)
    g2 = random_multigraph(32, n, 1600, 0.1, 0.05, 30, False)
    mp = random_mapping(33, n, unmapped, bijective)
    o = ob.Oracle(n, g1, g2)
    ref = o.first_improvement(mp)
    with engine_for(n, g1, g2, mp) as e:
        acc, st = e.first_improvement(verify=True)
        np.testing.assert_array_equal(acc, ref["accepted"])
        np.testing.assert_array_equal(e.mapping(), ref["mapping"])
        assert st["final_score"] == ref["final_score"] and st["sweeps"] == ref["sweeps"]
        assert st["rows_cleared"] > 0 and st["rows"] == st["sweeps"] * n
        if not bijective:
            assert e.sweep_config()[0] != 2  # colliding images: not the sparse kernel


def test_row_best_pass_matches_oracle():
    inst = bmatch.Instance.powerlaw(6, 150, 2000)
    o = ob.Oracle(inst.n, inst.g1, inst.g2)
    m_ref, acc_ref = o.row_best_pass(inst.p0)
    with engine_for(inst.n, inst.g1, inst.g2, inst.p0) as e:
        acc, st = e.row_best(max_sweeps=1)
        np.testing.assert_array_equal(acc, acc_ref)
        np.testing.assert_array_equal(e.mapping(), m_ref)


def test_global_best_steps_match_oracle():
    inst = bmatch.Instance.powerlaw(8, 80, 900)
    o = ob.Oracle(inst.n, inst.g1, inst.g2)
    cur = inst.p0.copy()
    want = []
    for _ in range(15):
        r = o.all_pairs(cur)
        a, b, d = r["best"]
        if d <= 0:
            break
        want.append((a, b, d))
        cur[a], cur[b] = cur[b], cur[a]
    with engine_for(inst.n, inst.g1, inst.g2, inst.p0) as e:
        acc, st = e.global_best(max_steps=15)
        np.testing.assert_array_equal(acc[: len(want)], np.array(want, np.int32).reshape(-1, 3))
        np.testing.assert_array_equal(e.mapping(), cur)


@pytest.mark.parametrize("cols", [1, 2, 4])
@pytest.mark.parametrize("case", [CASES[1], CASES[6], CASES[7]], ids=["dups_loops", "negative", "wide"])
def test_all_pairs_every_slab_width(case, cols, monkeypatch):
    # the autotuner may pick any width; each one must give the oracle's matrix
    monkeypatch.setenv("BM_GAIN_COLS", str(cols))
    n, g1, g2, mp = make_case(case, seed=17)
    ref = ob.Oracle(n, g1, g2).all_pairs(mp, matrix=True)
    with engine_for(n, g1, g2, mp) as e:
        got = e.sweep_all_pairs(rows=True, matrix=True)
    np.testing.assert_array_equal(got["matrix"], ref["matrix"])
    assert got["best"] == ref["best"]


@pytest.mark.parametrize("cols", [2, 4])
def test_all_pairs_long_lists_packed_flush(cols, monkeypatch):
    # weights up to 4000 make the packed uint16 sums widen every 16 entries;
    # lists of ~200 entries cross that interval many times
    monkeypatch.setenv("BM_GAIN_COLS", str(cols))
    n = 60
    g1 = random_multigraph(77, n, 6000, 0.05, 0.02, 4000)
    g2 = random_multigraph(78, n, 6000, 0.05, 0.02, 4000)
    mp = random_mapping(79, n)
    ref = ob.Oracle(n, g1, g2).all_pairs(mp, matrix=True)
    with engine_for(n, g1, g2, mp) as e:
        got = e.sweep_all_pairs(rows=True, matrix=True)
    np.testing.assert_array_equal(got["matrix"], ref["matrix"])
    assert got["best"] == ref["best"]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_all_pairs_sparse_kernel(case, monkeypatch):
    # the sparse gain kernel (taken when p is injective on mapped vertices and
    # G2 never names vertex 0; the other cases fall back to the dense kernel)
    monkeypatch.setenv("BM_SWEEP_KERNEL", "sparse")
    n, g1, g2, mp = make_case(case, seed=23)
    ref = ob.Oracle(n, g1, g2).all_pairs(mp, matrix=True)
    with engine_for(n, g1, g2, mp) as e:
        got = e.sweep_all_pairs(rows=True, matrix=True)
    np.testing.assert_array_equal(got["matrix"], ref["matrix"])
    assert got["best"] == ref["best"]
    np.testing.assert_array_equal(got["row_best_b"][1:n], ref["row_best_b"][1:n])


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_all_pairs_sparse_kernel_split_rows(case, monkeypatch):
    # rows longer than BM_SPARSE_SPLIT comb entries are spread over several
    # blocks that add their partial rows into X with atomics
    monkeypatch.setenv("BM_SWEEP_KERNEL", "sparse")
    monkeypatch.setenv("BM_SPARSE_SPLIT", "8")
    monkeypatch.setenv("BM_SPARSE_DENSE_LEN", "6")
    n, g1, g2, mp = make_case(case, seed=29)
    ref = ob.Oracle(n, g1, g2).all_pairs(mp, matrix=True)
    with engine_for(n, g1, g2, mp) as e:
        for _ in range(2):  # the split rows are zeroed again on every sweep
            got = e.sweep_all_pairs(rows=True, matrix=True)
            np.testing.assert_array_equal(got["matrix"], ref["matrix"])
            assert got["best"] == ref["best"]


@pytest.mark.parametrize("dense_len", ["100000", "20", "0"])
def test_sparse_split_hub_rows_after_moves(dense_len, monkeypatch):
    # split rows, and G2 lists longer than BM_SPARSE_DENSE_LEN streamed as whole Q rows
    monkeypatch.setenv("BM_SWEEP_KERNEL", "sparse")
    monkeypatch.setenv("BM_SPARSE_SPLIT", "100")
    monkeypatch.setenv("BM_SPARSE_DENSE_LEN", dense_len)
    inst = bmatch.Instance.powerlaw(37, 400, 9000)  # hubs far above 100 entries
    o = ob.Oracle(inst.n, inst.g1, inst.g2)
    cur = inst.p0.copy()
    with engine_for(inst.n, inst.g1, inst.g2, inst.p0) as e:
        for _ in range(4):
            got = e.sweep_all_pairs(rows=True)
            ref = o.all_pairs(cur)
            assert got["best"] == ref["best"]
            np.testing.assert_array_equal(got["row_best_delta"][1:inst.n], ref["row_best_delta"][1:inst.n])
            a, b, d = ref["best"]
            e.apply_swap(a, b)
            cur[a], cur[b] = cur[b], cur[a]


@pytest.mark.parametrize("kernel", ["sparse", "dense"])
def test_global_best_after_moves_each_kernel(kernel, monkeypatch):
    # p^-1 must follow every applied swap, rotation and set permutation
    monkeypatch.setenv("BM_SWEEP_KERNEL", kernel)
    inst = bmatch.Instance.powerlaw(31, 90, 1100)
    o = ob.Oracle(inst.n, inst.g1, inst.g2)
    cur = inst.p0.copy()
    with engine_for(inst.n, inst.g1, inst.g2, inst.p0) as e:
        e.apply_rot3(4, 9, 17, 3)
        cur = ob.apply_g3(cur, 4, 9, 17, 3)
        e.apply_perm([5, 6, 7, 8], [2, 0, 3, 1])
        cur = ob.apply_perm(cur, [5, 6, 7, 8], [2, 0, 3, 1])
        for _ in range(10):
            got = e.sweep_all_pairs(rows=True)
            ref = o.all_pairs(cur)
            assert got["best"] == ref["best"]
            np.testing.assert_array_equal(got["row_best_delta"][1:inst.n], ref["row_best_delta"][1:inst.n])
            a, b, d = ref["best"]
            if d <= 0:
                break
            e.apply_swap(a, b)
            cur[a], cur[b] = cur[b], cur[a]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_row_kernel_after_moves(case):
    # the row kernels (sparse where p is injective, dense otherwise) and the
    # incrementally maintained X(v,v) after swaps, rotations and set moves
    n, g1, g2, mp = make_case(case, seed=29)
    o = ob.Oracle(n, g1, g2)
    rng = np.random.default_rng(5)
    cur = mp.copy()
    with engine_for(n, g1, g2, mp) as e:
        e.eval_swap_row(1, 1, n)  # X(v,v) valid from here on
        for step in range(12):
            a, b, c3 = (int(x) for x in rng.choice(np.arange(1, n + 1), 3, replace=False))
            if step % 3 == 0:
                e.apply_swap(a, b)
                cur[a], cur[b] = cur[b], cur[a]
            elif step % 3 == 1:
                e.apply_rot3(a, b, c3, 4)
                cur = ob.apply_g3(cur, a, b, c3, 4)
            else:
                e.apply_perm([a, b, c3], [1, 2, 0])
                cur = ob.apply_perm(cur, [a, b, c3], [1, 2, 0])
            row = int(rng.integers(1, n + 1))
            ds = {bb: o.delta_b(cur, row, bb) for bb in range(1, n + 1) if bb != row}
            pos = [bb for bb in sorted(ds) if ds[bb] > 0]
            assert e.eval_swap_row(row, 1, n) == ((pos[0], ds[pos[0]]) if pos else (0, 0))
            bb = max(ds, key=lambda x: (ds[x], -x))
            assert e.eval_swap_row(row, 1, n, bmatch.BM_ROW_ARGMAX) == (bb, ds[bb])


@pytest.mark.parametrize("case", ["plain", "loops_dups", "vertex0", "nonbijective", "negative"])
def test_apply_swap_scored_is_the_exact_change(case):
    n = 70
    neg = case == "negative"
    g1 = random_multigraph(81, n, 700, 0.0 if case == "plain" else 0.15, 0.0 if case == "plain" else 0.05,
                           zero_vertex=case == "vertex0", negative=neg)
    g2 = random_multigraph(82, n, 700, 0.1, 0.05)
    mp = random_mapping(83, n, 0.1, bijective=case != "nonbijective")
    o = ob.Oracle(n, g1, g2)
    rng = np.random.default_rng(84)
    cur = mp.copy()
    with bmatch.Engine(0) as e:
        e.upload(n, g1, g2)
        e.set_mapping(mp)
        s = e.score()
        for _ in range(40):
            a, b = (int(x) for x in rng.choice(np.arange(0 if case == "vertex0" else 1, n + 1), 2, replace=False))
            want = o.delta_exact(cur, a, b)
            got = e.apply_swap_scored(a, b)
            cur[a], cur[b] = cur[b], cur[a]
            assert got == want, (a, b)
            s += got
            assert s == e.score() == o.score_opt(cur)


def dense_all_pairs(o, n, mp):
    """calculateSwapDeltasKernel's matrix, row bests (b > a) and global best from the oracle's Delta_G2."""
    dm = np.zeros((n + 1, n + 1), np.int32)
    for a in range(1, n + 1):
        for b in range(a + 1, n + 1):
            dm[a, b] = dm[b, a] = o.delta_g2(mp, a, b)
    rb = np.zeros(n + 1, np.int32)
    rd = np.full(n + 1, np.iinfo(np.int32).min, np.int32)
    best = (0, 0, 0)
    for a in range(1, n):
        row = dm[a, a + 1:]
        j = int(np.argmax(row))
        rb[a], rd[a] = a + 1 + j, row[j]
        if row[j] > best[2]:
            best = (a, a + 1 + j, int(row[j]))
    return dm, rb, rd, best


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_all_pairs_dense_matches_oracle(case):
    n, g1, g2, mp = make_case(case, seed=23)
    o = ob.Oracle(n, g1, g2)
    dm, rb, rd, best = dense_all_pairs(o, n, mp)
    with engine_for(n, g1, g2, mp) as e:
        got = e.sweep_all_pairs(rows=True, matrix=True, kind=bmatch.BM_DELTA_DENSE)
        lists = e.sweep_all_pairs(rows=True, matrix=True)
    np.testing.assert_array_equal(got["matrix"], dm)
    np.testing.assert_array_equal(got["row_best_b"][1:n], rb[1:n])
    np.testing.assert_array_equal(got["row_best_delta"][1:n], rd[1:n])
    assert got["best"] == best
    if case[0] in ("dups_loops", "vertex0", "dense", "negative"):
        assert (got["matrix"] != lists["matrix"]).any()  # the case separates the two deltas


@pytest.mark.parametrize("verify", [False, True])
def test_global_best_dense_steps_match_oracle(verify):
    n, g1, g2, mp = make_case(CASES[1], seed=29)  # repeated lines and self-loops
    o = ob.Oracle(n, g1, g2)
    cur = mp.copy()
    want = []
    for _ in range(12):
        _, _, _, (a, b, d) = dense_all_pairs(o, n, cur)
        if d <= 0:
            break
        want.append((a, b, d))
        cur[a], cur[b] = cur[b], cur[a]
    with engine_for(n, g1, g2, mp) as e:
        acc, st = e.global_best(max_steps=12, verify=verify, kind=bmatch.BM_DELTA_DENSE)
        np.testing.assert_array_equal(acc, np.array(want, np.int32).reshape(-1, 3))
        np.testing.assert_array_equal(e.mapping(), cur)
        exact = o.score_opt(cur)
    if verify:  # the exact score, and every step whose kernel delta was not the score change counted
        starts = [apply_swaps(mp, want[:i]) for i in range(len(want))]
        changes = [o.delta_exact(m, a, b) for m, (a, b, _) in zip(starts, want)]
        assert st["final_score"] == exact == st["initial_score"] + sum(changes)
        assert st["mismatches"] == sum(d != ch for (_, _, d), ch in zip(want, changes))
    else:  # the reference's own bookkeeping: score += delta
        assert st["final_score"] == st["initial_score"] + sum(d for _, _, d in want)


def apply_swaps(mp, swaps):
    m = mp.copy()
    for a, b, _ in swaps:
        m[a], m[b] = m[b], m[a]
    return m
