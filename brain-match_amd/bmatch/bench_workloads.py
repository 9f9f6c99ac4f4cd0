"""Measurements of the configurations other than the all-pairs headline.

Each returns one dict in bench.py's JSON shape.  Workloads (seeded, synthetic):
  plain     power-law pair n=1,000 / 20,000 lines, random start, pair-swap
            first-improvement (graph-alignment-optimized.c:513-581) to
            convergence on the GPU
  plain-c2  the n=18,524 / 3.0M pair, first-improvement over the first rows
  rot3      the n=1,000 pair, 2^24 two-hop triples per launch, both rotations
  permset   the n=1,000 pair, 4,096 8-vertex sets x 40,320 permutations
"""
import time

import numpy as np

import bmatch


def _base(metric, value, unit, steps, ms, config, higher=True, dtype="int32"):
    return {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": steps, "warmup": 0,
            "ms_per_step": ms, "higher_is_better": higher, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
            "data": "synthetic (seeded)", "config": config}


def plain(args):
    inst = bmatch.Instance.powerlaw(1, 1000, 20000)
    with bmatch.Engine(0) as e:
        e.upload(inst.n, inst.g1, inst.g2)
        e.set_mapping(inst.p0)
        s0 = e.score()
        t = time.perf_counter()
        acc, st = e.first_improvement(verify=False)
        el = time.perf_counter() - t
    out = _base("plain first-improvement: exact score gain per second", (st["final_score"] - s0) / el,
                "score/s", int(st["sweeps"]), el * 1000 / max(st["sweeps"], 1),
                {"workload": "power-law pair n=1000 E=20000, random start, first-improvement to convergence",
                 "n": inst.n})
    out.update({"initial_score": s0, "final_score": int(st["final_score"]), "accepted": int(st["accepted"]),
                "row_searches": int(st["evals_rows"]), "seconds": el})
    return out


def plain_c2(args):
    inst = bmatch.Instance.chunglu(2, 18524, 3_000_000)
    rows = 400
    with bmatch.Engine(0) as e:
        e.upload(inst.n, inst.g1, inst.g2)
        e.set_mapping(inst.p0)
        s0 = e.score()
        e.eval_swap_row(1, 1, inst.n)
        t = time.perf_counter()
        acc, st = e.first_improvement(verify=False, max_rows=rows)
        st["rows"] = min(st["rows"], rows)
        el = time.perf_counter() - t
    out = _base("plain first-improvement rows/s at n=18.5k", st["rows"] / el, "rows/s", int(st["rows"]),
                el * 1000 / st["rows"],
                {"workload": f"Chung-Lu pair n=18524 E=3.0M, first {rows} rows of a first-improvement sweep",
                 "n": inst.n})
    out.update({"accepted": int(st["accepted"]), "score_gain": int(st["final_score"] - s0), "seconds": el,
                "evals_per_s": st["rows"] * (inst.n - 1) / el})
    return out


def rot3(args):
    inst = bmatch.Instance.powerlaw(1, 1000, 20000)
    count = 1 << 24
    tri = inst.triples_2hop(7, count)
    with bmatch.Engine(0) as e:
        e.upload(inst.n, inst.g1, inst.g2)
        e.set_mapping(inst.p0)
        e.moves_set(3, tri)
        e.moves_eval(bmatch.BM_ROT3_ROTATIONS)
        e.set_timing(True)
        e.reset_timing()
        steps = max(args.steps // 4, 3)
        t = time.perf_counter()
        for _ in range(steps):
            best = e.moves_eval(bmatch.BM_ROT3_ROTATIONS)
        el = (time.perf_counter() - t) / steps
        n_k, ms_k = e.kernel_time("rot3")
    out = _base("3-rotation deltas per second", 2 * count / el, "rotation evals/s", steps, el * 1000,
                {"workload": "power-law pair n=1000 E=20000, 2^24 two-hop triples per launch, both rotations",
                 "n": inst.n, "triples": count})
    out.update({"best": best, "kernel_avg_ms": ms_k / max(n_k, 1)})
    return out


def permset(args):
    inst = bmatch.Instance.powerlaw(1, 1000, 20000)
    sets = bmatch.gen_sets(9, inst.n, 8, 4096)
    with bmatch.Engine(0) as e:
        e.upload(inst.n, inst.g1, inst.g2)
        e.set_mapping(inst.p0)
        e.moves_set(8, sets)
        e.moves_eval(0)
        e.set_timing(True)
        e.reset_timing()
        steps = max(args.steps // 2, 3)
        t = time.perf_counter()
        for _ in range(steps):
            best = e.moves_eval(0)
        el = (time.perf_counter() - t) / steps
        n_k, ms_k = e.kernel_time("permset")
    out = _base("8-set permutation deltas per second", 4096 * 40320 / el, "permutation evals/s", steps, el * 1000,
                {"workload": "power-law pair n=1000 E=20000, 4096 8-vertex sets x 40320 permutations per launch",
                 "n": inst.n})
    out.update({"best": best, "kernel_avg_ms": ms_k / max(n_k, 1)})
    return out


def run(name, args):
    return {"plain": plain, "plain-c2": plain_c2, "rot3": rot3, "permset": permset}[name](args)
