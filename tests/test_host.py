"""Host layer without a GPU: ABI exports, file formats, generators, CLI argument handling."""
import os
import re
import subprocess

import numpy as np
import pytest

import bmatch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in ("bmatch.h", "bm_host.h"):
        with open(os.path.join(ROOT, "include", h)) as f:
            text = re.sub(r"/\*.*?\*/", "", f.read(), flags=re.S)
        names |= set(re.findall(r"\b(bm_[a-z0-9_]+)\s*\(", text))
    return {n for n in names if not n.endswith("_hook")}


def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", bmatch.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (bm_\w+)", out))
    missing = declared_symbols() - exported
    assert not missing, missing
    assert bmatch.lib().bm_abi_version() == 1


def test_error_strings():
    L = bmatch.lib()
    for code in range(0, -8, -1):
        assert L.bm_strerror(code)


def test_no_device_is_reported_not_crashed():
    # a device index that cannot exist must come back as an error code
    L = bmatch.lib()
    import ctypes as C

    h = C.c_void_p()
    rc = L.bm_create(4096, C.byref(h))
    assert rc == -5 and not h.value


def test_powerlaw_generator_invariants():
    a = bmatch.Instance.powerlaw(1, 1000, 20000)
    b = bmatch.Instance.powerlaw(1, 1000, 20000)
    c = bmatch.Instance.powerlaw(2, 1000, 20000)
    for x, y in zip(a.g1 + a.g2, b.g1 + b.g2):
        np.testing.assert_array_equal(x, y)  # pure function of the seed
    assert not np.array_equal(a.g1[1], c.g1[1])
    s, d, w = a.g1
    assert len(s) == 20000 and len(a.g2[0]) == 20000
    assert (s != d).all() and len(set(zip(s.tolist(), d.tolist()))) == 20000
    assert w.min() >= 1 and w.max() <= 255
    assert sorted(a.pi[1:].tolist()) == list(range(1, 1001)) and sorted(a.p0[1:].tolist()) == list(range(1, 1001))
    # G2 keeps 90% of pi(G1)
    e1 = set(zip(a.pi[s].tolist(), a.pi[d].tolist()))
    e2 = set(zip(a.g2[0].tolist(), a.g2[1].tolist()))
    assert 18000 <= len(e1 & e2) <= 18100  # fresh lines may re-draw a deleted pair


def test_chunglu_generator_invariants():
    a = bmatch.Instance.chunglu(3, 2000, 60000)
    s, d, w = a.g1
    assert len(s) == 60000 and (s != d).all() and len(set(zip(s.tolist(), d.tolist()))) == 60000
    assert w.min() >= 1 and w.max() <= 2000
    assert 4.0 < w.mean() < 8.0  # round(LogNormal(1, 1.2)) has mean about 5.6
    moved = (a.p0[1:] != a.pi[1:]).mean()
    assert 0.2 < moved <= 0.3


def test_sets_and_triples():
    sets = bmatch.gen_sets(5, 100, 8, 500)
    assert sets.min() >= 1 and sets.max() <= 100
    assert all(len(set(r.tolist())) == 8 for r in sets)
    inst = bmatch.Instance.powerlaw(4, 300, 3000)
    tri = inst.triples_2hop(9, 2000)
    nb = {}
    for x, y in zip(inst.g1[0].tolist(), inst.g1[1].tolist()):
        nb.setdefault(x, set()).add(y)
        nb.setdefault(y, set()).add(x)
    for u, v, w in tri.tolist():
        assert len({u, v, w}) == 3
        two = set(nb[u]) | {z for x in nb[u] for z in nb[x]}
        assert v in two and w in two


def test_edge_csv_reader(tmp_path):
    p = tmp_path / "g.csv"
    p.write_text("Source,Target,Weight\n1,2,5\nbad line\n3,4,7\n2,1,-3\n\n9,9,1\n")
    s, d, w, mx, bad = bmatch.read_edges_csv(str(p))
    assert s.tolist() == [1, 3, 2, 9] and d.tolist() == [2, 4, 1, 9] and w.tolist() == [5, 7, -3, 1]
    assert mx == 9 and bad == 2


def test_mapping_readers(tmp_path):
    p = tmp_path / "m.csv"
    p.write_text("Male,Female\nm12,f34\n5,6\n12,f7\nx3,y4\n")
    # load_benchmark_mapping: "%d,%d" or optional m/f prefixes
    a, b = bmatch.read_pairs(str(p), "optimized")
    assert a.tolist() == [12, 5, 12, 0] and b.tolist() == [34, 6, 7, 0]
    # load_mapping (permset): first character always dropped
    a, b = bmatch.read_pairs(str(p), "permset")
    assert a.tolist() == [12, 0, 2, 3] and b.tolist() == [34, 0, 7, 4]
    # load_initial_matching: "%d,%d" only
    a, b = bmatch.read_pairs(str(p), "plain")
    assert a.tolist() == [5] and b.tolist() == [6]


def test_readers_on_empty_and_header_only_files(tmp_path):
    for text in ("", "Source,Target,Weight\n", "Source,Target,Weight"):
        p = tmp_path / "g.csv"
        p.write_text(text)
        s, d, w, mx, bad = bmatch.read_edges_csv(str(p))
        assert len(s) == len(d) == len(w) == 0 and mx == -1 and bad == 0
        for fmt in ("optimized", "permset", "plain"):
            a, b = bmatch.read_pairs(str(p), fmt)
            assert len(a) == len(b) == 0
    # a last line without its newline is read like the others
    p = tmp_path / "g.csv"
    p.write_text("h\n1,2,3\n4,5,6")
    s, d, w, mx, bad = bmatch.read_edges_csv(str(p))
    assert s.tolist() == [1, 4] and d.tolist() == [2, 5] and w.tolist() == [3, 6] and bad == 0


def test_mapping_writer_round_trip(tmp_path):
    p = tmp_path / "out.csv"
    mp = np.array([0, 3, 0, 1, 2], np.int32)
    bmatch.write_mapping_mf(str(p), mp)
    assert p.read_text() == "Male Node ID,Female Node ID\nm1,f3\nm3,f1\nm4,f2\n"
    a, b = bmatch.read_pairs(str(p), "optimized")
    back = np.zeros(5, np.int32)
    back[a] = b
    np.testing.assert_array_equal(back, mp)


def _run(cli, *args, cwd=None):
    return subprocess.run([os.path.join(bmatch.BIN_DIR, cli), *args], capture_output=True, text=True, cwd=cwd,
                          timeout=60)


def test_cli_usage_errors():
    r = _run("graph-alignment-optimized", "a", "b")
    assert r.returncode == 1 and "Usage:" in r.stderr
    r = _run("graph-alignment-permset-parallel", "a")
    assert r.returncode == 1 and "Usage:" in r.stderr
    r = _run("refine-matching-hillclimb", "a", "b", "c")
    assert r.returncode == 1 and "Usage:" in r.stderr


def test_cli_missing_inputs(tmp_path):
    r = _run("graph-alignment-optimized", *(str(tmp_path / x) for x in "abcdef"))
    assert r.returncode == 1 and "Failed to load male graph" in r.stderr
    r = _run("refine-matching-hillclimb", *(str(tmp_path / x) for x in "abcd"))
    assert r.returncode == 1 and "Cannot open graph file" in r.stderr
    r = _run("graph-alignment-permset-parallel", *(str(tmp_path / x) for x in "abc"))
    assert r.returncode == 1 and "Error loading input files" in r.stderr


def test_bm_gen_writes_reference_formats(tmp_path):
    r = _run("bm-gen", "powerlaw", "7", "50", "300", str(tmp_path))
    assert r.returncode == 0, r.stderr
    s, d, w, mx, bad = bmatch.read_edges_csv(str(tmp_path / "male.csv"))
    assert len(s) == 300 and bad == 0 and mx <= 50
    a, b = bmatch.read_pairs(str(tmp_path / "init_mf.csv"), "optimized")
    assert sorted(a.tolist()) == list(range(1, 51)) and sorted(b.tolist()) == list(range(1, 51))
    assert (tmp_path / "order_male.csv").read_text().startswith("node,order\n")
