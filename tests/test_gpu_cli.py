"""The drop-in CLIs on the GPU, against the reference programs' recorded behaviour."""
import json
import os
import re
import subprocess

import numpy as np
import pytest

import bmatch
import oracle_binding as ob

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def run(cli, *args, cwd=None, env=None, timeout=300):
    e = dict(os.environ)
    e.update(env or {})
    return subprocess.run([os.path.join(bmatch.BIN_DIR, cli), *args], capture_output=True, text=True, cwd=cwd,
                          env=e, timeout=timeout)


def read_mf(path, n):
    mp = np.zeros(n + 1, np.int32)
    with open(path) as f:
        next(f)
        for line in f:
            a, b = line.strip().split(",")
            mp[int(a.lstrip("m"))] = int(b.lstrip("f"))
    return mp


@pytest.mark.parametrize("name", sorted(d for d in os.listdir(GOLD) if d.startswith("ga_opt_")))
def test_graph_alignment_optimized_replays_reference(name, tmp_path):
    d = os.path.join(GOLD, name)
    with open(os.path.join(d, "expected.json")) as f:
        exp = json.load(f)
    for o in ("om.csv", "of.csv"):
        (tmp_path / o).write_text("node,order\n")
    out = tmp_path / "out.csv"
    r = run("graph-alignment-optimized", os.path.join(d, "male.csv"), os.path.join(d, "female.csv"),
            str(tmp_path / "om.csv"), str(tmp_path / "of.csv"), os.path.join(d, "init.csv"), str(out),
            env={"BM_ROWS_PER_SWEEP": str(exp["rows_per_sweep"])})
    assert r.returncode == 0, r.stderr[-2000:]
    # progress samples, parsed the same way as from the reference's stdout
    rows = [int(x.replace(",", "")) for x in re.findall(r"Processed: ([\d,]+)/", r.stdout)]
    scores = [int(x.replace(",", "")) for x in re.findall(r"Current score: ([\d,]+)", r.stdout)]
    imps = [int(x.replace(",", "")) for x in re.findall(r"Improvements found: ([\d,]+)", r.stdout)]
    got = list(zip(rows, scores, imps))
    changes = exp["log_changes"]
    want, j = [], 0
    for row in range(exp["log_every"], rows[-1] + 1, exp["log_every"]):
        while j + 1 < len(changes) and changes[j + 1][0] <= row:
            j += 1
        want.append((row, changes[j][1], changes[j][2]))
    assert got == want
    np.testing.assert_array_equal(read_mf(out, exp.get("vertex_range", exp["n"])),
                                  np.array(exp["final_mapping"], np.int32))
    assert f"Final score: {changes[-1][1]:,}" in r.stdout
    if "initial_score" in exp:
        assert f"Initial alignment score: {exp['initial_score']:,}" in r.stdout
        # the per-swap lines of :570-576, in order
        mm = [(int(a), int(b)) for a, b in
              re.findall(r"Score mismatch after swap - calculated: (-?\d+), verified: (-?\d+)", r.stderr)]
        assert mm == [tuple(x) for x in exp["mismatches"]]


def test_refine_matching_hillclimb_cli(tmp_path):
    inst = bmatch.Instance.powerlaw(21, 150, 2000)
    # ids shifted so that index order differs from id order
    off = 1000
    with open(tmp_path / "g1.csv", "w") as f:
        f.write("src,dst,w\n")
        for a, b, w in zip(*inst.g1):
            f.write(f"{a + off},{b + off},{w}\n")
    with open(tmp_path / "g2.csv", "w") as f:
        f.write("src,dst,w\n")
        for a, b, w in zip(*inst.g2):
            f.write(f"{a + 5},{b + 5},{w}\n")
    with open(tmp_path / "m.csv", "w") as f:
        f.write("a,b\n")
        for i in range(1, inst.n + 1):
            f.write(f"{i + off},{inst.p0[i] + 5}\n")
    r = run("refine-matching-hillclimb", str(tmp_path / "g1.csv"), str(tmp_path / "g2.csv"), str(tmp_path / "m.csv"),
            str(tmp_path / "out.csv"))
    assert r.returncode == 0, r.stderr[-2000:]
    init = int(re.search(r"Initial matching score: (\d+)", r.stdout).group(1))
    final = int(re.search(r"Final matching score: (\d+)", r.stdout).group(1))
    o = ob.Oracle(inst.n, inst.g1, inst.g2)
    assert init == o.score_opt(inst.p0)
    lines = (tmp_path / "out.csv").read_text().splitlines()
    assert lines[0] == "Male ID,Female ID" and len(lines) == inst.n + 1
    mp = np.zeros(inst.n + 1, np.int32)
    for ln in lines[1:]:
        a, b = map(int, ln.split(","))
        mp[a - off] = b - 5
    assert sorted(mp[1:].tolist()) == list(range(1, inst.n + 1))
    assert final == o.score_opt(mp) > init
    # converged: no single swap improves the written matching
    ref = o.first_improvement(mp, max_sweeps=1)
    assert len(ref["accepted"]) == 0


def test_permset_cli_writes_improvements(tmp_path):
    inst = bmatch.Instance.powerlaw(22, 200, 2500)
    r = run("bm-gen", "powerlaw", "22", "200", "2500", str(tmp_path))
    assert r.returncode == 0
    r = run("graph-alignment-permset-parallel", "male.csv", "female.csv", "init_mf.csv", cwd=str(tmp_path),
            env={"BM_MAX_ITERS": "30", "BM_NUM_GROUPS": "512", "BM_GROUP_SIZE": "4", "BM_SEED": "5"})
    assert r.returncode == 0, r.stderr[-2000:]
    outs = sorted(p for p in os.listdir(tmp_path) if p.startswith("improved_mapping_score_"))
    assert outs, r.stdout[-2000:]
    scores = [int(re.match(r"improved_mapping_score_(\d+)_", p).group(1)) for p in outs]
    o = ob.Oracle(inst.n, inst.g1, inst.g2)
    best = max(scores)
    mp = np.zeros(inst.n + 1, np.int32)
    a, b = bmatch.read_pairs(str(tmp_path / [p for p in outs if f"_{best}_" in p][0]), "permset")
    mp[a] = b
    assert o.score_opt(mp) == best > o.score_opt(inst.p0)
