"""Write the golden fixtures of tests/golden/ from the reference programs.

Runs ONLY where the reference is mounted: it executes the reference's own
C programs, compiled unmodified from /root/reference/src into oracle/_ref/
by `make -C oracle ref`.  The fixtures it writes are plain data (input CSVs,
the program's printed numbers, its output mapping) and are committed, so the
tests never need the reference.

  score_matching.json : score-matching (src/score-matching.c) on small pairs
                        with repeated lines, self-loops, m/f prefixes and
                        unmapped vertices.
  ga_opt_<name>/       : graph-alignment-optimized (src/graph-alignment-
                        optimized.c) run until a full 18,524-row sweep
                        accepts nothing: its progress log (rows, score,
                        improvements every 20 rows, :519-534) and the mapping
                        file it last wrote (:536-540).

  ga_opt_multi/        : the same on a pair with lines at vertex 0, self-loops
                        and repeated lines, also recording the per-swap
                        "Score mismatch after swap" lines (:570-576).

Usage: python tests/golden/make_golden.py [score] [n100] [n240] [multi]
"""
import json
import os
import re
import signal
import subprocess
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref")
sys.path.insert(0, os.path.join(ROOT, "brain-match_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

NUM_NODES = 18524  # graph-alignment-optimized.c:15


def write_edges(path, src, dst, w, header="Source,Target,Weight"):
    with open(path, "w") as f:
        f.write(header + "\n")
        for a, b, c in zip(src, dst, w):
            f.write(f"{a},{b},{c}\n")


def write_map_mf(path, mp):
    with open(path, "w") as f:
        f.write("Male Node ID,Female Node ID\n")
        for i in range(1, len(mp)):
            if mp[i] != 0:
                f.write(f"m{i},f{mp[i]}\n")


def score_matching_cases(outdir):
    from conftest import random_multigraph

    cases = []
    rng = np.random.default_rng(2024)
    specs = [("plain", 30, 200, 0.0, 0.0, 0.0), ("dups", 30, 250, 0.3, 0.0, 0.0),
             ("loops", 25, 220, 0.1, 0.1, 0.0), ("unmapped", 40, 300, 0.1, 0.05, 0.25)]
    tmp = os.path.join(outdir, "_tmp_sm")
    os.makedirs(tmp, exist_ok=True)
    for name, n, m, dup, loops, unm in specs:
        g1 = random_multigraph(int(rng.integers(1 << 30)), n, m, dup, loops, 50)
        g2 = random_multigraph(int(rng.integers(1 << 30)), n, m, dup, loops, 50)
        mp = np.zeros(n + 1, np.int32)
        mp[1:] = rng.permutation(n) + 1
        mp[1:][rng.random(n) < unm] = 0
        for fmt in ("mf", "plain"):
            p1, p2, pm = (os.path.join(tmp, x) for x in ("g1.csv", "g2.csv", "m.csv"))
            write_edges(p1, *g1)
            write_edges(p2, *g2)
            with open(pm, "w") as f:
                f.write("Male,Female\n")
                for i in range(1, n + 1):
                    if mp[i]:
                        f.write(f"m{i},f{mp[i]}\n" if fmt == "mf" else f"{i},{mp[i]}\n")
            out = subprocess.run([os.path.join(REF, "score-matching"), p1, p2, pm], capture_output=True,
                                 text=True, check=True, timeout=60).stdout.strip().splitlines()[-1]
            cases.append({"name": f"{name}_{fmt}", "n": n, "g1": [a.tolist() for a in g1],
                          "g2": [a.tolist() for a in g2], "mapping": mp.tolist(), "map_format": fmt,
                          "score": int(out)})
    for x in ("g1.csv", "g2.csv", "m.csv"):
        os.remove(os.path.join(tmp, x))
    os.rmdir(tmp)
    with open(os.path.join(outdir, "score_matching.json"), "w") as f:
        json.dump({"program": "src/score-matching.c", "cases": cases}, f, separators=(",", ":"))
    print("score-matching:", [(c["name"], c["score"]) for c in cases])


_NUM = re.compile(r"([\d,]+)")


_MISMATCH = re.compile(r"Score mismatch after swap - calculated: (-?\d+), verified: (-?\d+)")
_INITIAL = re.compile(r"Initial alignment score: (-?[\d,]+)")


def run_ga_opt(name, inst_seed, n, m, outdir, budget_s=400, graphs=None):
    """graphs: (g1, g2, p0, description) instead of the seeded power-law pair."""
    import bmatch

    if graphs is None:
        inst = bmatch.Instance.powerlaw(inst_seed, n, m)
        g1, g2, p0, desc = inst.g1, inst.g2, inst.p0, f"bmatch.Instance.powerlaw({inst_seed}, {n}, {m})"
    else:
        g1, g2, p0, desc = graphs
    d = os.path.join(outdir, f"ga_opt_{name}")
    os.makedirs(d, exist_ok=True)
    write_edges(os.path.join(d, "male.csv"), *g1)
    write_edges(os.path.join(d, "female.csv"), *g2)
    write_map_mf(os.path.join(d, "init.csv"), p0)
    work = os.path.join(d, "_run")
    os.makedirs(work, exist_ok=True)
    for o in ("order_m.csv", "order_f.csv"):
        with open(os.path.join(work, o), "w") as f:
            f.write("node,order\n")
    outmap = os.path.join(work, "out.csv")
    cmd = [os.path.join(REF, "graph-alignment-optimized"), os.path.join(d, "male.csv"),
           os.path.join(d, "female.csv"), os.path.join(work, "order_m.csv"), os.path.join(work, "order_f.csv"),
           os.path.join(d, "init.csv"), outmap]
    proc = subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, cwd=work)
    log, cur = [], {}
    mismatches, initial = [], None
    t0 = time.time()
    converged_at = None
    for line in proc.stdout:
        mm = _MISMATCH.search(line)
        if mm:
            mismatches.append((int(mm.group(1)), int(mm.group(2))))
            continue
        mi = _INITIAL.search(line)
        if mi and initial is None:
            initial = int(mi.group(1).replace(",", ""))
        if "Processed:" in line:
            cur = {"rows": int(line.split("Processed:")[1].split("/")[0].strip().replace(",", ""))}
        elif "Current score:" in line and "rows" in cur:
            cur["score"] = int(line.split(":")[1].strip().replace(",", ""))
        elif "Improvements found:" in line and "score" in cur:
            cur["impr"] = int(line.split(":")[1].strip().replace(",", ""))
            log.append((cur["rows"], cur["score"], cur["impr"]))
            cur = {}
            # converged: a whole sweep of rows without a new improvement
            rows, _, impr = log[-1]
            if converged_at is None:
                back = [e for e in log if e[0] <= rows - NUM_NODES - 40]
                if back and back[-1][2] == impr and rows > 2 * NUM_NODES:
                    converged_at = rows
            elif rows >= converged_at + 200:
                break
        if time.time() - t0 > budget_s:
            break
    proc.send_signal(signal.SIGTERM)
    proc.wait(timeout=30)
    if converged_at is None:
        raise RuntimeError(f"{name}: not converged within {budget_s}s")
    # the converged mapping (rewritten every 25 rows, identical once converged)
    with open(outmap) as f:
        lines = f.read().splitlines()
    pairs = []
    for ln in lines[1:]:
        a, b = ln.split(",")
        pairs.append((int(a[1:]), int(b[1:])))
    # with lines at G2's vertex 0 a swap with one of the empty rows above n
    # (the program's vertices run to NUM_NODES) can pay, so the mapping file
    # may name vertices past n; every image 1..n is still there once
    vrange = NUM_NODES if graphs is not None else n
    final = np.zeros(vrange + 1, np.int32)
    for a, b in pairs:
        final[a] = b
    if sorted(final[final > 0].tolist()) != list(range(1, n + 1)):
        raise RuntimeError(f"{name}: mapping file incomplete (caught mid-write?)")
    # keep the log entries where something changed, plus the last one
    compact = [log[0]] + [log[i] for i in range(1, len(log)) if log[i][1:] != log[i - 1][1:]] + [log[-1]]
    with open(os.path.join(d, "expected.json"), "w") as f:
        json.dump({"program": "src/graph-alignment-optimized.c", "n": n, "m": m, "generator": desc,
                   "rows_per_sweep": NUM_NODES, "log_every": 20, "log_changes": compact,
                   "log_last_row": log[-1][0], "final_mapping": final.tolist(),
                   "converged_at_row": converged_at, "initial_score": initial, "vertex_range": vrange,
                   "mismatches": mismatches}, f, separators=(",", ":"))
    for x in os.listdir(work):
        os.remove(os.path.join(work, x))
    os.rmdir(work)
    print(f"ga_opt {name}: n={n} m={m} entries={len(log)} changes={len(compact)} final score={log[-1][1]} "
          f"improvements={log[-1][2]} initial={initial} mismatches={len(mismatches)} ({time.time() - t0:.0f}s)")


def multigraph_pair(seed, n, m):
    """A pair with every input the reference's score and delta treat apart:
    lines leaving and entering vertex 0 in both graphs (the score loop starts
    at vertex 1, graph-alignment-optimized.c:295, while the delta counts a
    line from 0 as an in-entry of its target, and W2's row 0 is read for
    every line from 0), self-loops, and repeated lines with fresh weights
    (the dense cell takes the last, :185)."""
    from conftest import random_multigraph

    rng = np.random.default_rng(seed)
    g = []
    for k in range(2):
        s, d, w = random_multigraph(seed + 1 + k, n, m, dup_frac=0.1, selfloop_frac=0.03, wmax=30)
        # a few lines at vertex 0 (no 0 -> 0), placed at random positions in the file
        zs = rng.integers(1, n + 1, 12)
        zsrc = np.where(np.arange(12) < 8, 0, zs).astype(np.int32)
        zdst = np.where(np.arange(12) < 8, zs, 0).astype(np.int32)
        zw = rng.integers(1, 31, 12).astype(np.int32)
        s, d, w = np.concatenate([s, zsrc]), np.concatenate([d, zdst]), np.concatenate([w, zw])
        perm = rng.permutation(len(s))
        g.append((s[perm], d[perm], w[perm]))
    p0 = np.zeros(n + 1, np.int32)
    p0[1:] = rng.permutation(n) + 1
    return g[0], g[1], p0, (f"tests/golden/make_golden.py multigraph_pair({seed}, {n}, {m}): random_multigraph "
                            f"with dup_frac=0.1, selfloop_frac=0.03, wmax=30, plus 8 lines from and 4 into vertex 0 "
                            f"per graph")


def main():
    if not os.path.isdir("/root/reference"):
        sys.exit("the reference is not mounted here; fixtures are committed under tests/golden/")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    only = sys.argv[1:]
    if not only or "score" in only:
        score_matching_cases(HERE)
    if not only or "n100" in only:
        run_ga_opt("n100", 11, 100, 700, HERE)
    if not only or "n240" in only:
        run_ga_opt("n240", 12, 240, 2400, HERE)
    if not only or "multi" in only:
        run_ga_opt("multi", 0, 90, 700, HERE, graphs=multigraph_pair(31, 90, 700))


if __name__ == "__main__":
    main()
