#!/usr/bin/env python3
"""Benchmark of the MI355X move-delta engine on the full-size sweep workload.

One step = one all-pairs Delta_B sweep (graph-alignment-optimized.c:321-408
for every pair 1 <= a < b <= n, with per-row and global arg-max) over the
synthetic n = 18,524 / 3.0M-line graph pair, everything resident in HBM.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
                    [--cpu-seconds S | --no-cpu-baseline] [--sweep-only]
                    [--calibrate-copy MiB] [--calibrate-gather MiB] [--calibrate-stream MiB]
                    [--workload NAME] [--window S] [--jobs J]

Prints one JSON line:
  value      Delta_B evaluations per second of the sweep;
  roofline   the dominant kernel (the all-pairs gain kernel): its average
             launch time from HIP events on the engine's stream, the bytes it
             moves between L2 and memory per launch (rocprofv3 counters of the
             same kernel and workload, stored under profiles/) over that time
             against the 8 TB/s HBM peak, and beside it the algorithmic bytes
             of the per-evaluation model (12 B per list term + 16 B);
  served     the served stream (Poisson arrivals of refinement jobs): jobs/s
             with p50/p99 time-to-result;
  batched_vs_plain  exact score gain per second of the batched refinement
             against plain first-improvement on the same pair and start;
  cpu_baseline  the reference's own OpenMP first-improvement program
             (graph-alignment-optimized-parallel, built from the reference
             sources into oracle/_ref/) on this host's cores, from a mapping
             no swap improves, so that its time goes to delta evaluations;
             without that program, the oracle's restatement of the
             alignment-omp.c:598-622 row-best loop.
--sweep-only skips the served, batched and CPU legs (profiler runs).
--workload selects the other measurements (bmatch/bench_workloads.py).
"""
import argparse
import json
import os
import re
import signal
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "brain-match_amd"))

METRIC = "swap-delta evals/s & HBM GB/s at n=18.5k; refine jobs/s + p99 latency"
HBM_PEAK_GBS = 8000.0
N, E, SEED = 18524, 3_000_000, 2
COUNTERS = os.path.join(ROOT, "profiles", "gain_allpairs_n18524_counters.json")
REF_PARALLEL = os.path.join(ROOT, "oracle", "_ref", "graph-alignment-optimized-parallel")
REF_NUM_NODES = 18524  # graph-alignment-optimized-parallel.c:16


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(n, two_e):
    """Sum over pairs a<b of 12*T(a,b) + 16, T = list entries of a and b.

    sum_{a<b} (d(a) + d(b)) = (n - 1) * sum_v d(v) = (n - 1) * 2E.
    """
    return 12 * (n - 1) * two_e + 16 * n * (n - 1) // 2


def host_cpu():
    """(logical CPUs usable, physical cores among them, threads to use, CPU model)."""
    cpus = sorted(os.sched_getaffinity(0))
    cores = set()
    for c in cpus:
        t = f"/sys/devices/system/cpu/cpu{c}/topology/"
        try:
            with open(t + "physical_package_id") as f1, open(t + "core_id") as f2:
                cores.add((f1.read().strip(), f2.read().strip()))
        except OSError:
            cores.add(("?", str(c)))
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    # one GPU's share of the host: the OpenMP thread count the environment sets, at most the physical cores
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(cores)
    return len(cpus), len(cores), min(len(cores), share), model


def cpu_baseline_reference(inst, mapping, seconds):
    """Times graph-alignment-optimized-parallel (built unmodified from the reference
    sources) on this pair: its rows/s between progress lines (every 100 rows,
    graph-alignment-optimized-parallel.c:521-533), each row evaluating m2 = 1..18,524
    (m2 != m1) with the OpenMP loop of :544-585."""
    import bmatch

    logical, physical, threads, model = host_cpu()
    work = tempfile.mkdtemp(prefix="bm_cpu_", dir="/tmp")
    files = {k: os.path.join(work, k + ".csv") for k in ("male", "female", "om", "of", "init", "out")}
    bmatch.write_edges_csv(files["male"], inst.g1)
    bmatch.write_edges_csv(files["female"], inst.g2)
    bmatch.write_mapping_mf(files["init"], mapping)
    for k in ("om", "of"):
        with open(files[k], "w") as f:
            f.write("node,order\n")
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
    # a pseudo-terminal keeps the program's stdout line-buffered, so each progress line is seen when printed
    import pty

    master, slave = pty.openpty()
    proc = subprocess.Popen([REF_PARALLEL] + [files[k] for k in ("male", "female", "om", "of", "init", "out")],
                            stdout=slave, stderr=subprocess.DEVNULL, stdin=subprocess.DEVNULL, cwd=work, env=env,
                            start_new_session=True)
    os.close(slave)
    pat = re.compile(rb"Processed: ([\d,]+)/")
    first = last = None
    t_start = time.perf_counter()
    buf = b""
    try:
        while True:
            try:
                chunk = os.read(master, 65536)
            except OSError:
                break
            if not chunk:
                break
            now = time.perf_counter()
            buf += chunk
            *lines, buf = buf.split(b"\n")
            for line in lines:
                m = pat.search(line)
                if m:
                    rows = int(m.group(1).replace(b",", b""))
                    if rows >= 200 and first is None:
                        first = (rows, now)
                    elif first is not None:
                        last = (rows, now)
            if last and last[1] - first[1] >= seconds:
                break
            if now - t_start > seconds + 120:
                break
    finally:
        os.killpg(proc.pid, signal.SIGKILL)
        proc.wait()
        os.close(master)
        for p in files.values():
            if os.path.exists(p):
                os.remove(p)
        os.rmdir(work)
    if not first or not last:
        return None
    rows_s = (last[0] - first[0]) / (last[1] - first[1])
    evals_s = rows_s * (REF_NUM_NODES - 1)
    return {"value": evals_s, "unit": "evals/s", "cores": threads, "kind": "reference",
            "sample": f"rows {first[0]}..{last[0]} of graph-alignment-optimized-parallel (src/graph-alignment-"
                      f"optimized-parallel.c, built from the reference sources) on this pair, {last[1] - first[1]:.1f} s, "
                      f"from a mapping no swap improves (every row is {REF_NUM_NODES - 1} Delta_B evaluations, "
                      f"graph-alignment-optimized.c:321-408, and the mapping dump every 25 rows)",
            "rows_per_s": rows_s, "logical_cpus": logical, "physical_cores": physical, "cpu_model": model,
            "omp": f"OMP_NUM_THREADS={threads} OMP_PROC_BIND=close OMP_PLACES=cores"}


def cpu_baseline_port(inst, seconds):
    """The oracle's restatement of alignment-omp.c:598-622 (OpenMP over m2, critical arg-max) on the
    reference layout (EdgeMap lists, short** W2), rows timed with the mapping held fixed."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_binding as ob

    logical, physical, threads, model = host_cpu()
    os.environ.setdefault("OMP_PROC_BIND", "close")
    o = ob.Oracle(inst.n, inst.g1, inst.g2)
    evals, rows, t0 = 0, 0, time.perf_counter()
    step = 16
    while time.perf_counter() - t0 < seconds:
        ev, _ = o.omp_rowbest_rows(inst.p0, 1 + rows, 1 + rows + step, threads)
        evals += ev
        rows += step
    el = time.perf_counter() - t0
    return {"value": evals / el, "unit": "evals/s", "cores": threads, "kind": "port",
            "sample": f"rows 1..{rows} of the n={inst.n} row-best sweep ({evals} Delta_B evaluations, {el:.1f} s): "
                      f"oracle restatement of alignment-omp.c:598-622 on the reference layout",
            "logical_cpus": logical, "physical_cores": physical, "cpu_model": model,
            "omp": f"OMP_NUM_THREADS={threads} OMP_PROC_BIND=close"}


def calibrate_copy(mib):
    """A plain copy kernel of known size, for the memory counters."""
    import torch

    x = torch.ones(mib * (1 << 20) // 4, dtype=torch.int32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"calibrate_copy_bytes": mib * (1 << 20), "launches": 3}))


def calibrate_stream(mib, reps=10):
    """The read and read+write rates that torch's own streaming kernels (a
    float sum, a copy) reach over a mib-MiB buffer, timed with events: library
    kernels beside the 8 TB/s peak."""
    import torch

    x = torch.ones(mib * (1 << 20) // 4, dtype=torch.int32, device="cuda")
    y = torch.empty_like(x)
    xf = x.view(torch.float32)
    out = {"calibrate_stream_bytes": mib * (1 << 20), "reps": reps}
    for name, fn, moved in (("read_gbs", lambda: xf.sum(), 1), ("copy_gbs", lambda: y.copy_(x), 2)):
        for _ in range(3):
            fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        out[name] = moved * mib * (1 << 20) / (ms * 1e-3) / 1e9
        out[name.replace("_gbs", "_ms")] = ms
    print(json.dumps(out))


def calibrate_gather(mib, count_log2=24):
    """Reference gathers for the memory read counters: count random 4-byte words
    from a mib-MiB buffer (a miss per word), then the same count of consecutive
    words; the difference of their read bytes is the cost of the misses."""
    import torch

    x = torch.ones(mib * (1 << 20) // 4, dtype=torch.int32, device="cuda")
    g = torch.Generator(device="cuda")
    g.manual_seed(1)
    k = 1 << count_log2
    rnd = torch.randint(0, x.numel(), (k,), device="cuda", generator=g)
    seq = torch.arange(k, device="cuda")
    for idx in (rnd, seq, rnd, seq):
        y = torch.index_select(x, 0, idx)
    torch.cuda.synchronize()
    lines = int(torch.unique(rnd // 32).numel())  # distinct 128-byte lines of the random gathers
    print(json.dumps({"calibrate_gather_words": k, "buffer_bytes": mib * (1 << 20), "distinct_128B_lines": lines,
                      "order": ["random", "sequential", "random", "sequential"], "out_sum": int(y.sum())}))


def roofline(kname, kernel_ms, alg):
    out = {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None, "traffic": None,
           "kernel": kname, "kernel_avg_ms": kernel_ms, "algorithmic_bytes_per_launch": alg,
           "algorithmic_gbs": alg / (kernel_ms / 1000.0) / 1e9}
    if os.path.exists(COUNTERS):
        with open(COUNTERS) as f:
            pmc = json.load(f)
        if pmc.get("kernel") == kname:
            traffic = pmc["read_bytes_per_launch"] + pmc["write_bytes_per_launch"]
            gbs = traffic / (kernel_ms / 1000.0) / 1e9
            out.update({"achieved": gbs, "frac": gbs / HBM_PEAK_GBS, "traffic": traffic,
                        "traffic_source": f"profiles/{os.path.basename(COUNTERS)}: rocprofv3 --pmc "
                                          "TCC_EA0_RDREQ_DRAM_32B_sum + TCC_EA0_WRREQ_WRITE_DRAM_32B_sum (x 32 B) of "
                                          f"this kernel on this workload (kernel source {pmc.get('source', '?')})",
                        "limiters": pmc.get("limiters")})
        lim = pmc.get("limiters") or {}
        if lim:
            out["limited_by"] = (
                "no unit saturated: the traffic past L2 (MALL hits included, so an upper bound on HBM bytes) at "
                f"{gbs / HBM_PEAK_GBS:.0%} of the HBM peak, LDS active {lim.get('lds_active_frac', 0):.0%} of the "
                f"time ({lim.get('lds_bank_conflict_over_active', 0):.0%} of that bank conflicts from the "
                f"scattered row adds), VALU about {lim.get('valu_busy_est', 0):.0%} busy, L2 hit rate "
                f"{lim.get('l2_hit_rate', 0):.0%}; the per-word list walk (load, min, LDS add) sets the time")
    return out


def served_leg(jobs, rate):
    from bmatch import bench_workloads

    class A:
        pass

    a = A()
    a.jobs, a.rate = jobs, rate
    r = bench_workloads.serve(a)
    return {k: r[k] for k in ("value", "unit", "p50_s", "p99_s", "mean_queue_s", "mean_prepare_s",
                              "mean_service_s", "score_gain_per_s")} | {"jobs": jobs, "arrival_rate": rate,
                                                                         "workload": r["config"]["workload"]}


def config_legs(steps):
    """Short runs of the n=1,000 configurations: plain first-improvement to convergence, 2^24 two-hop
    3-rotations per launch, 4,096 8-sets x 40,320 permutations per launch, and the batched 8-set step
    with its exact score check."""
    from bmatch import bench_workloads

    class A:
        pass

    a = A()
    a.steps = steps
    keep = ("value", "unit", "ms_per_step", "steps")
    out = {}
    r = bench_workloads.plain(a)
    out["c1_plain"] = {k: r[k] for k in keep} | {k: r[k] for k in ("seconds", "accepted", "initial_score",
                                                                   "final_score")}
    r = bench_workloads.rot3(a)
    out["c3_rot3"] = {k: r[k] for k in keep} | {"form": r["form"], "kernel_avg_ms": r["kernel_avg_ms"],
                                                "best": r["best"]}
    r = bench_workloads.permset(a)
    out["c4_permset"] = {k: r[k] for k in keep} | {"kernel_avg_ms": r["kernel_avg_ms"], "best": r["best"]}
    r = bench_workloads.batched_permset(a)
    out["c4_batched_permset"] = {k: r[k] for k in keep} | {"sets_applied": r["sets_applied"],
                                                           "score_gain": r["score_gain"], "verified": True}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--sweep-only", action="store_true", help="only the sweep (profiler runs)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MiB")
    ap.add_argument("--calibrate-gather", type=int, default=0, metavar="MiB")
    ap.add_argument("--calibrate-stream", type=int, default=0, metavar="MiB")
    ap.add_argument("--window", type=float, default=0.0, help="seconds per leg of the batched comparison")
    ap.add_argument("--jobs", type=int, default=0, help="jobs of the serve workload")
    ap.add_argument("--rate", type=float, default=20.0, help="arrivals per second of the serve workload")
    ap.add_argument("--workload", default="allpairs",
                    choices=["allpairs", "plain", "plain-c2", "rot3", "permset", "batched", "batched-c1",
                             "batched-permset", "batched-many", "quick", "q3", "triples", "serve", "serve-sweep"],
                    help="allpairs (default, the headline); the others measure the other configurations")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the sweep would shard rows "
              "across GPUs with one 8-byte arg-max all-reduce per step; not built).", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.calibrate_gather:
        calibrate_gather(args.calibrate_gather)
        return
    if args.calibrate_stream:
        calibrate_stream(args.calibrate_stream)
        return

    import bmatch
    from bmatch import bench_workloads

    if args.workload != "allpairs":
        print(json.dumps(bench_workloads.run(args.workload, args)))
        return

    log(f"generating n={N} E={E} (seed {SEED})")
    inst = bmatch.Instance.chunglu(SEED, N, E)
    eng = bmatch.Engine(0)
    eng.upload(inst.n, inst.g1, inst.g2)
    eng.set_mapping(inst.p0)
    n = inst.n
    pairs = n * (n - 1) // 2
    two_e = 2 * len(inst.g1[0])
    for _ in range(args.warmup):
        eng.sweep_all_pairs()
    log(f"warmup done, timing {args.steps} sweeps")
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
        "roofline": roofline(kname, gain_avg_ms, algorithmic_bytes(n, two_e)),
    }
    converged = None
    if not args.sweep_only:
        # batched refinement against plain first-improvement, same pair and start, on this engine; the
        # batched run's end mapping (no improving swap left) is the CPU baseline's start
        secs = args.window or 5.0
        prm = bmatch.refine_params(pairs=1, triples=1024, sets=256, set_k=4, rounds=0, steps_per_check=8, seed=1)
        log("batched vs plain")
        eng.refine_configure(prm)
        b = bench_workloads._refine_window(eng, secs, chunk=8)
        converged = eng.mapping()
        conv_best = eng.sweep_all_pairs()["best"][2]
        # the drop-in first-improvement sweep on that mapping (every row searched, none accepts when
        # no swap improves): the GPU side of the CPU baseline below, which runs on the same mapping
        log("plain sweep from the converged mapping")
        t = time.perf_counter()
        _, st_c = eng.first_improvement(verify=False, max_sweeps=1)
        el_c = time.perf_counter() - t
        out["plain_converged_sweep"] = {"rows_per_s": st_c["rows"] / el_c,
                                        "evals_per_s": st_c["rows"] * (n - 1) / el_c, "rows": st_c["rows"],
                                        "accepted": st_c["accepted"], "seconds": el_c,
                                        "row_searches": st_c["evals_rows"]}
        eng.set_mapping(inst.p0)
        p = bench_workloads._plain_window(eng, max(b["seconds"], 1.0))
        out["batched_vs_plain"] = {
            "ratio": b["gain_per_s"] / p["gain_per_s"] if p["gain_per_s"] > 0 else None,
            "batched_gain_per_s": b["gain_per_s"], "plain_gain_per_s": p["gain_per_s"],
            "batched": {k: b[k] for k in ("seconds", "steps", "moves", "initial_score", "final_score")},
            "plain": {k: p[k] for k in ("seconds", "rows", "moves", "initial_score", "final_score")},
            "converged": conv_best <= 0}
    eng.close()
    if not args.sweep_only:
        log("served stream")
        out["served"] = served_leg(args.jobs or 40, args.rate)
        log("n=1,000 configurations (plain, 3-rotations, 8-sets, batched 8-sets)")
        out["configs_1000"] = config_legs(args.steps)
    if args.no_cpu_baseline or args.sweep_only:
        out["cpu_baseline"] = None
    else:
        log(f"CPU baseline (~{args.cpu_seconds:.0f} s)")
        cb = None
        if os.path.exists(REF_PARALLEL) and n == REF_NUM_NODES and converged is not None:
            cb = cpu_baseline_reference(inst, converged, args.cpu_seconds)
        out["cpu_baseline"] = cb or cpu_baseline_port(inst, args.cpu_seconds)
        out["cpu_baseline"]["effective_gbs"] = out["cpu_baseline"]["value"] * algorithmic_bytes(n, two_e) / pairs / 1e9
    print(json.dumps(out))


if __name__ == "__main__":
    main()
