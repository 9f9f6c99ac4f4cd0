"""CPU-baseline calibration: the oracle's restatement of the reference OpenMP
row-best loop against the reference's own programs, on the same input.

TEST INFRASTRUCTURE: runs the reference programs built from their sources
into oracle/_ref/ (make -C oracle ref) and the oracle library; nothing of the
product imports it.

On the full-size pair (n = 18,524, 3.0M lines, the bench instance), from a
start mapping (default: the generator's hidden permutation), it times
  - graph-alignment-optimized-parallel: rows/s between its progress lines
    (every 100 rows, graph-alignment-optimized-parallel.c:521-533); each row
    is 18,523 Delta_B evaluations;
  - alignment-omp: rows processed between its first progress lines (every
    4,000 rows, alignment-omp.c:575-588); row m1 is 18,524 - m1 evaluations;
  - the oracle's orc_omp_rowbest_rows (alignment-omp.c:598-622 restated on
    the reference layout) over the same rows with the same thread count;
and prints evaluations per second for each with the ratio port / program.
Threads: OMP_NUM_THREADS (default: the physical cores), OMP_PROC_BIND=close.

    python oracle/calibrate_cpu.py [--mapping path.csv] [--seconds S]
"""
import argparse
import json
import os
import pty
import re
import signal
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "brain-match_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))
NUM_NODES = 18524


def physical_cores():
    cores = set()
    for c in sorted(os.sched_getaffinity(0)):
        t = f"/sys/devices/system/cpu/cpu{c}/topology/"
        try:
            with open(t + "physical_package_id") as f1, open(t + "core_id") as f2:
                cores.add((f1.read().strip(), f2.read().strip()))
        except OSError:
            cores.add(("?", str(c)))
    return len(cores)


def run_program(exe, files, threads, pattern, want_rows, seconds):
    """Rows/s of a reference program between progress samples (timestamps taken as lines arrive)."""
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
    master, slave = pty.openpty()
    proc = subprocess.Popen([exe] + files, stdout=slave, stderr=subprocess.DEVNULL, stdin=subprocess.DEVNULL,
                            env=env, cwd=os.path.dirname(files[0]), start_new_session=True)
    os.close(slave)
    pat = re.compile(pattern)
    samples, buf, t0 = [], b"", time.perf_counter()
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
                    samples.append((int(m.group(1).replace(b",", b"")), now))
            if len(samples) >= want_rows or (samples and now - samples[0][1] > seconds) or now - t0 > 4 * seconds + 600:
                break
    finally:
        os.killpg(proc.pid, signal.SIGKILL)
        proc.wait()
        os.close(master)
    return samples


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mapping", default="", help="start mapping (m<id>,f<id> CSV); default: the hidden permutation")
    ap.add_argument("--seconds", type=float, default=60.0)
    args = ap.parse_args()
    import bmatch
    import oracle_binding as ob

    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or physical_cores()
    inst = bmatch.Instance.chunglu(2, NUM_NODES, 3_000_000)
    if args.mapping:
        pairs = bmatch.read_pairs(args.mapping, fmt="optimized")
        start = pairs["mapping"] if isinstance(pairs, dict) else pairs
    else:
        start = inst.pi
    work = tempfile.mkdtemp(prefix="bm_cal_", dir="/tmp")
    f = {k: os.path.join(work, k + ".csv") for k in ("male", "female", "om", "of", "init", "out")}
    bmatch.write_edges_csv(f["male"], inst.g1)
    bmatch.write_edges_csv(f["female"], inst.g2)
    bmatch.write_mapping_mf(f["init"], start)
    for k in ("om", "of"):
        with open(f[k], "w") as fh:
            fh.write("node,order\n")
    files = [f[k] for k in ("male", "female", "om", "of", "init", "out")]
    out = {"threads": threads, "physical_cores": physical_cores(), "logical_cpus": len(os.sched_getaffinity(0)),
           "omp": f"OMP_NUM_THREADS={threads} OMP_PROC_BIND=close OMP_PLACES=cores",
           "start": args.mapping or "hidden permutation"}
    with open("/proc/cpuinfo") as fh:
        out["cpu_model"] = next((l.split(":", 1)[1].strip() for l in fh if l.startswith("model name")), "unknown")
    o = ob.Oracle(inst.n, inst.g1, inst.g2)
    # graph-alignment-optimized-parallel: rows 200 .. 200 + ~seconds of rows
    s = run_program(os.path.join(HERE, "_ref", "graph-alignment-optimized-parallel"), files, threads,
                    rb"Processed: ([\d,]+)/", 10 ** 9, args.seconds)
    s = [x for x in s if x[0] >= 200]
    par = (s[-1][0] - s[0][0]) * (NUM_NODES - 1) / (s[-1][1] - s[0][1])
    out["parallel"] = {"rows": [s[0][0], s[-1][0]], "evals_per_s": par}
    # the port over rows of the same count, mapping fixed (every row 18,523 evaluations in -parallel; the
    # row-best loop evaluates m2 > m1 only, so compare per evaluation)
    t = time.perf_counter()
    ev, _ = o.omp_rowbest_rows(start, 1, 1 + (s[-1][0] - s[0][0]), threads)
    port = ev / (time.perf_counter() - t)
    out["port"] = {"rows": [1, s[-1][0] - s[0][0]], "evals_per_s": port}
    out["port_over_parallel"] = port / par
    # alignment-omp: its first two progress lines (rows 4,000 and 8,000)
    s = run_program(os.path.join(HERE, "_ref", "alignment-omp"), files, threads, rb"Processed: ([\d,]+)/", 2,
                    10 * args.seconds)
    if len(s) >= 2:
        ev_omp = sum(NUM_NODES - m1 for m1 in range(s[0][0] + 1, s[1][0] + 1))
        omp = ev_omp / (s[1][1] - s[0][1])
        out["alignment_omp"] = {"rows": [s[0][0], s[1][0]], "evals_per_s": omp}
        t = time.perf_counter()
        ev, _ = o.omp_rowbest_rows(start, s[0][0] + 1, s[1][0] + 1, threads)
        out["port_same_rows"] = {"rows": [s[0][0] + 1, s[1][0]], "evals_per_s": ev / (time.perf_counter() - t)}
        out["port_over_alignment_omp"] = out["port_same_rows"]["evals_per_s"] / omp
    for p in f.values():
        if os.path.exists(p):
            os.remove(p)
    os.rmdir(work)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
