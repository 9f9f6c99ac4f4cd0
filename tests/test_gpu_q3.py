"""The changed-vertex triple search of q3.cu on the GPU (bm_drive_q3), step
for step against the oracle's restatement: same seed, same counters, same
best mapping and scores."""
import numpy as np
import pytest

import bmatch
import oracle_binding as ob
from conftest import random_mapping, random_multigraph

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", ["powerlaw", "multigraph", "vertex0"])
@pytest.mark.parametrize("seed", [1, 2])
def test_q3_matches_oracle(case, seed):
    if case == "powerlaw":
        inst = bmatch.Instance.powerlaw(40 + seed, 150, 1800)
        n, g1, g2, mp = inst.n, inst.g1, inst.g2, inst.p0
    else:
        n = 70
        zero = case == "vertex0"
        g1 = random_multigraph(50 + seed, n, 900, 0.15, 0.05, 30, zero_vertex=zero)
        g2 = random_multigraph(60 + seed, n, 900, 0.15, 0.05, 30, zero_vertex=zero)
        mp = random_mapping(70 + seed, n)
    o = ob.Oracle(n, g1, g2)
    want_map, want = o.q3(mp, seed=seed, max_evals=80, best_window=30)
    with bmatch.Engine(0) as e:
        e.upload(n, g1, g2)
        e.set_mapping(mp)
        got_map, got = e.q3(seed=seed, max_evals=80, best_window=30)
        np.testing.assert_array_equal(e.mapping(), got_map)
        assert e.score() == got["best_score"]
    assert got == want
    np.testing.assert_array_equal(got_map, want_map)
    assert got["moves"] > 0 and got["best_score"] >= got["initial_score"]
