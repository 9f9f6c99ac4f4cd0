"""Device buffers freed by one context are handed to the next one (the
library's block cache): results must not depend on what a reused buffer held.

Pairs of different sizes are refined and swept one after another, in
contexts created and destroyed in between and side by side; each result
equals the first run of the same pair in a fresh sequence, and the sweep's
best move equals the oracle's Delta_B of that move.
"""
import numpy as np
import pytest

import bmatch
import oracle_binding as ob

pytestmark = pytest.mark.gpu


def run_pair(inst, prm):
    with bmatch.Engine(0) as e:
        e.upload(inst.n, inst.g1, inst.g2)
        e.set_mapping(inst.p0)
        best = e.sweep_all_pairs()["best"]
        e.refine_configure(prm)
        st = e.refine_run(max_steps=40, patience=4)
        return best, st["final_score"], st["moves"], e.mapping().copy()


def test_results_do_not_depend_on_reused_buffers():
    insts = [bmatch.Instance.powerlaw(90 + i, n, 12 * n) for i, n in enumerate((300, 120, 420, 200))]
    prm = bmatch.refine_params(pairs=1, triples=500, sets=40, set_k=4, rounds=8, steps_per_check=2, seed=3)
    first = [run_pair(inst, prm) for inst in insts]
    for inst, (best, _, _, _) in zip(insts, first):
        a, b, d = best
        assert d == ob.Oracle(inst.n, inst.g1, inst.g2).delta_b(inst.p0, a, b)
    # again in the opposite order, so each context takes blocks another size left behind
    again = [run_pair(inst, prm) for inst in reversed(insts)][::-1]
    for x, y in zip(first, again):
        assert x[:3] == y[:3]
        np.testing.assert_array_equal(x[3], y[3])
    # two contexts alive at once, each with blocks freed by the other sizes
    with bmatch.Engine(0) as e1, bmatch.Engine(0) as e2:
        for e, inst in ((e1, insts[1]), (e2, insts[2])):
            e.upload(inst.n, inst.g1, inst.g2)
            e.set_mapping(inst.p0)
            e.refine_configure(prm)
        s1, s2 = e1.refine_run(max_steps=40, patience=4), e2.refine_run(max_steps=40, patience=4)
        assert (s1["final_score"], s1["moves"]) == first[1][1:3]
        assert (s2["final_score"], s2["moves"]) == first[2][1:3]
