#!/usr/bin/env python3
"""Benchmark of the MI355X move-delta engine on the full-size sweep workload.

One step = one all-pairs Delta_B sweep (graph-alignment-optimized.c:321-408
for every pair 1 <= a < b <= n, with per-row and global arg-max) over the
synthetic n = 18,524 / 3.0M-edge graph pair, everything resident in HBM.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
                    [--cpu-seconds S | --no-cpu-baseline] [--calibrate-copy MiB]
                    [--workload NAME] [--window S] [--jobs J]

Prints one JSON line: metric/value (Delta_B evaluations per second), the
roofline of the dominant kernel (the all-pairs gain kernel the chooser
picked) with its average launch time measured by HIP events on the engine's
stream, and the CPU baseline (the oracle's restatement of the reference
OpenMP row-best loop, alignment-omp.c:598-622, timed on this host on a
bounded row sample).  --workload selects the other measurements: plain,
plain-c2, rot3, permset, batched, batched-c1, batched-permset, batched-many, serve
(bmatch/bench_workloads.py).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "brain-match_amd"))

METRIC = "swap-delta evals/s & HBM GB/s at n=18.5k; refine jobs/s + p99 latency"
HBM_PEAK_GBS = 8000.0
N, E, SEED = 18524, 3_000_000, 2
PMC_SUMMARY = os.path.join(ROOT, "profiles", "gain_allpairs_n18524_pmc.json")


def algorithmic_bytes(n, two_e):
    """Sum over pairs a<b of 12*T(a,b) + 16, T = list entries of a and b.

    sum_{a<b} (d(a) + d(b)) = (n - 1) * sum_v d(v) = (n - 1) * 2E.
    """
    return 12 * (n - 1) * two_e + 16 * n * (n - 1) // 2


def cpu_baseline(inst, seconds):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_binding as ob

    o = ob.Oracle(inst.n, inst.g1, inst.g2)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    evals, rows, t0 = 0, 0, time.perf_counter()
    step = 16
    while time.perf_counter() - t0 < seconds:
        ev, _ = o.omp_rowbest_rows(inst.p0, 1 + rows, 1 + rows + step, threads)
        evals += ev
        rows += step
    el = time.perf_counter() - t0
    return {"value": evals / el, "unit": "evals/s", "cores": threads, "kind": "port",
            "sample": f"rows 1..{rows} of the n={inst.n} row-best sweep ({evals} Delta_B evaluations, "
                      f"{el:.1f} s): oracle restatement of alignment-omp.c:598-622 (OpenMP over m2, critical "
                      f"arg-max) on the reference layout (EdgeMap lists, short** W2)"}


def calibrate_copy(mib):
    """A plain copy kernel of known size, for scaling FETCH_SIZE/WRITE_SIZE."""
    import torch

    x = torch.ones(mib * (1 << 20) // 4, dtype=torch.int32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"calibrate_copy_bytes": mib * (1 << 20), "launches": 3}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MiB")
    ap.add_argument("--window", type=float, default=0.0, help="seconds per leg of the batched comparison")
    ap.add_argument("--jobs", type=int, default=0, help="jobs of the serve workload")
    ap.add_argument("--workload", default="allpairs",
                    choices=["allpairs", "plain", "plain-c2", "rot3", "permset", "batched", "batched-c1",
                             "batched-permset", "batched-many", "quick", "serve"],
                    help="allpairs (default, the headline); the others measure the other configurations")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the sweep would shard rows "
              "across GPUs with one 8-byte arg-max all-reduce per step; not built).", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return

    import bmatch

    if args.workload != "allpairs":
        from bmatch import bench_workloads

        print(json.dumps(bench_workloads.run(args.workload, args)))
        return

    print(f"[bench] generating n={N} E={E} (seed {SEED})", file=sys.stderr, flush=True)
    inst = bmatch.Instance.chunglu(SEED, N, E)
    eng = bmatch.Engine(0)
    eng.upload(inst.n, inst.g1, inst.g2)
    eng.set_mapping(inst.p0)
    n = inst.n
    pairs = n * (n - 1) // 2
    two_e = 2 * len(inst.g1[0])
    for _ in range(args.warmup):
        eng.sweep_all_pairs()
    print(f"[bench] warmup done, timing {args.steps} sweeps", file=sys.stderr, flush=True)
    eng.set_timing(True)
    eng.reset_timing()
    t0 = time.perf_counter()
    best = None
    for _ in range(args.steps):
        best = eng.sweep_all_pairs()["best"]
    t1 = time.perf_counter()
    eng.set_timing(False)
    launches, gain_ms = eng.kernel_time("gain")
    kind, cols = eng.sweep_config()
    kname = "k_gain_sparse" if kind == 2 else f"k_gain<{cols}>"
    ms_step = (t1 - t0) * 1000.0 / args.steps
    gain_avg_ms = gain_ms / max(launches, 1)
    alg = algorithmic_bytes(n, two_e)
    achieved = alg / (gain_avg_ms / 1000.0) / 1e9
    traffic = None
    if os.path.exists(PMC_SUMMARY):
        with open(PMC_SUMMARY) as f:
            pmc = json.load(f)
        traffic = pmc.get("hbm_bytes_per_launch") if pmc.get("kernel") == kname else None
    out = {
        "metric": METRIC,
        "value": pairs / (ms_step / 1000.0),
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (seeded Chung-Lu pair; the reference's gm/gf are not distributed)",
        "config": {"workload": "full-size plain pairwise sweep: all-pairs Delta_B with row/global arg-max",
                   "n": n, "g1_edges": len(inst.g1[0]), "g2_edges": len(inst.g2[0]), "pairs_per_step": pairs,
                   "weights": "round(LogNormal(1.0,1.2)) clipped 1..2000", "init": "pi with 30% shuffled",
                   "parallelism": "single GPU"},
        "best_move": {"a": best[0], "b": best[1], "delta": best[2]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": kname, "kernel_avg_ms": gain_avg_ms, "algorithmic_bytes_per_launch": alg,
                     # the bytes the kernel really moves past L2 (PMC, profiles/), per second of this run
                     "traffic_gbs": traffic / (gain_avg_ms / 1000.0) / 1e9 if traffic else None,
                     "traffic_frac": traffic / (gain_avg_ms / 1000.0) / 1e9 / HBM_PEAK_GBS if traffic else None},
    }
    if not args.no_cpu_baseline:
        print(f"[bench] CPU baseline (~{args.cpu_seconds:.0f} s)", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(inst, args.cpu_seconds)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
