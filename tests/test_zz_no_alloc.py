"""The allocation-free pair-delta entry point (bm_eval_swaps_noalloc) on the GPU.

It must give bm_eval_swaps' deltas bit for bit, call after call (the second
call smaller than the first, so stale scratch contents would show), leave the
context's device footprint unchanged, and reject lists beyond its scratch.
"""
import numpy as np
import pytest

import bmatch
from conftest import random_mapping, random_multigraph

pytestmark = pytest.mark.gpu

MAX_PAIRS = 1 << 20


def pairs(seed, n, count):
    rng = np.random.default_rng(seed)
    a = rng.integers(0, n + 1, count).astype(np.int32)
    b = rng.integers(0, n + 1, count).astype(np.int32)
    b = np.where(a == b, (b + 1) % (n + 1), b).astype(np.int32)
    return a, b


@pytest.mark.parametrize("case", ["powerlaw", "multigraph"])
def test_noalloc_matches_eval_swaps(case):
    if case == "powerlaw":
        inst = bmatch.Instance.powerlaw(5, 3000, 40000)
        n, g1, g2, mp = inst.n, inst.g1, inst.g2, inst.p0
    else:  # self-loops, repeated lines, lines at vertex 0, unmapped vertices
        n = 400
        g1 = random_multigraph(8, n, 6000, 0.1, 0.03, 40, zero_vertex=True)
        g2 = random_multigraph(9, n, 6000, 0.1, 0.03, 40, zero_vertex=True)
        mp = random_mapping(10, n, 0.1)
    with bmatch.Engine(0) as e:
        e.upload(n, g1, g2)
        e.set_mapping(mp)
        before = e.device_bytes
        a1, b1 = pairs(1, n, MAX_PAIRS)  # the largest list the scratch holds
        got1 = e.eval_swaps_noalloc(a1, b1)
        a2, b2 = pairs(2, n, 5000)       # then a smaller one
        got2 = e.eval_swaps_noalloc(a2, b2)
        assert e.device_bytes == before
        e.apply_swap(5, 17)  # a moved mapping is seen
        a3, b3 = pairs(3, n, 777)
        got3 = e.eval_swaps_noalloc(a3, b3)
        assert e.device_bytes == before
        e.set_mapping(mp)
        want1 = e.eval_swaps(a1, b1)
        want2 = e.eval_swaps(a2, b2)
        e.apply_swap(5, 17)
        want3 = e.eval_swaps(a3, b3)
    np.testing.assert_array_equal(got1, want1)
    np.testing.assert_array_equal(got2, want2)
    np.testing.assert_array_equal(got3, want3)
    assert (want1 != 0).any()


def test_noalloc_rejects_oversize_and_bad_pairs():
    inst = bmatch.Instance.powerlaw(6, 500, 5000)
    with bmatch.Engine(0) as e:
        e.upload(inst.n, inst.g1, inst.g2)
        e.set_mapping(inst.p0)
        a, b = pairs(4, inst.n, MAX_PAIRS + 1)
        with pytest.raises(bmatch.BMError) as ex:
            e.eval_swaps_noalloc(a, b)
        assert ex.value.code == bmatch.BM_EINVAL and "scratch" in str(ex.value)
        with pytest.raises(bmatch.BMError):
            e.eval_swaps_noalloc(np.array([3], np.int32), np.array([3], np.int32))
        np.testing.assert_array_equal(e.eval_swaps_noalloc(a[:10], b[:10]), e.eval_swaps(a[:10], b[:10]))
