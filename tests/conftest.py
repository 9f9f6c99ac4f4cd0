import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "brain-match_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


def random_multigraph(seed, n, m, dup_frac=0.0, selfloop_frac=0.0, wmax=30, zero_vertex=False, negative=False):
    """Edge list in 'CSV order' with optional repeated pairs and self-loops.

    Repeated pairs get fresh weights, so the dense cell (last line wins)
    differs from the earlier lines, which stay in the per-vertex lists.
    """
    rng = np.random.default_rng(seed)
    lo = 0 if zero_vertex else 1
    src = rng.integers(lo, n + 1, m)
    dst = rng.integers(lo, n + 1, m)
    loops = rng.random(m) < selfloop_frac
    dst = np.where(loops, src, dst)
    same = (src == dst) & ~loops
    dst = np.where(same, (dst % n) + 1, dst)
    w = rng.integers(-wmax if negative else 1, wmax + 1, m)
    ndup = int(dup_frac * m)
    if ndup:
        idx = rng.integers(0, m, ndup)
        src = np.concatenate([src, src[idx]])
        dst = np.concatenate([dst, dst[idx]])
        w = np.concatenate([w, rng.integers(-wmax if negative else 1, wmax + 1, ndup)])
        perm = rng.permutation(len(src))
        src, dst, w = src[perm], dst[perm], w[perm]
    return (src.astype(np.int32), dst.astype(np.int32), w.astype(np.int32))


def random_mapping(seed, n, unmapped_frac=0.0, bijective=True):
    rng = np.random.default_rng(seed)
    m = np.zeros(n + 1, np.int32)
    if bijective:
        m[1:] = rng.permutation(n) + 1
    else:
        m[1:] = rng.integers(1, n + 1, n)
    if unmapped_frac:
        m[1:][rng.random(n) < unmapped_frac] = 0
    return m


@pytest.fixture(scope="session")
def engine_lib():
    import bmatch

    return bmatch.lib()
