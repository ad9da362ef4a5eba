"""Benchmark: reads partitioned per second on one MI355X.

Workload (the plain GPU path at medium size): 200 synthetic genomes of 1-5 Mb
with per-genome GC fraction U(0.3, 0.7) and ten 5 kb repeats each planted in
two genomes; 20,000,000 single-end reads of 150 bp with 1 % substitutions;
k = 31, m = 10, one scan, one segment, 16 chunks.

One step = the whole partitioning path over the FASTQ bytes already resident
in HBM: chunk index (host), record parse + 2-bit packing, m-mer histograms,
range plan, k-mer tuple enumeration, radix sort, frequency filter, union-find,
compress, pair step, partition plan and the device-side build of the 16
partitioned FASTQ streams.  Host file writes are not part of the step.

The same JSON line also reports:
  * "stages": per-kernel device time, algorithmic bytes and GB/s against the
    8 TB/s HBM peak (events on the library's stream);
  * "tools_wall": the three drop-in tools run end to end on the same reads
    (FASTQ in the page cache, output files on local disk);
  * "cpu_baseline": the reference implementation (built from its sources) on
    the first 2 M reads, all host cores up to 16, median of 3 runs;
  * "variants": the 3-minimizer variant on the same reads;
  * "batched": the 64-sample mix (sample i: own pool of 20-300 genomes of
    0.1-2 Mb seeded 100 + i, log-uniform 10k-5M reads of 100-150 bp, 0.5-2 %
    errors) resident in HBM, plain per-sample path vs one batch;
  * "served": 200 Poisson-arriving jobs (job j: seed 5000 + j, 20-100 genomes
    of 0.5-3 Mb, log-uniform 50k-2M reads of 150 bp) at 0.7 x the plain path's
    job rate, the request loop vs the plain per-job path (p50/p99 latency);
  * "out_of_core": the 100 M-read share (1000 genomes of 0.5-8 Mb,
    log-normal(1.5)), uploaded through pinned staging, scans from the budget.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--reads N] [--no-cpu-baseline] [--no-tools]
                       [--no-min3] [--no-batch] [--no-serve] [--no-c3]
"""
from __future__ import annotations

import argparse
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TUPLE_BYTES = 13  # u64 k-mer + u32 label + u8 adj


def measured_traffic(reads: int):
    """HBM bytes per radix pass from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary of this workload (counter runs cannot share a process with the
    timed run), or None when no summary exists for this read count."""
    p = os.path.join(ROOT, "profiles", "onesweep_kernel_traffic_c2_%dM.md" % (reads // 1_000_000))
    if reads % 1_000_000 or not os.path.exists(p):
        return None
    for line in open(p):
        if line.startswith("traffic_bytes_per_launch = "):
            return {"bytes": int(line.split("=")[1]), "source": os.path.relpath(p, ROOT)}
    return None


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads() -> int:
    """Host threads for the CPU baseline: the cores this process may use, at
    most 16 (the host share of one GPU on these nodes)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(sample_reads: int, K: int, threads: int, runs: int = 3):
    """The reference implementation (built from its sources into oracle/_ref)
    on the first sample_reads reads of the workload: genFastqIdx +
    genHistograms_metapart + metapart wall time with `threads` OpenMP threads,
    4 chunks per thread, one scan; the median of `runs` runs, and the per-stage
    line metapart prints (metapart.c:4100) of the median run."""
    import statistics
    import metapart_amd as mp
    ref = os.path.join(ROOT, "oracle", "_ref")
    if not os.path.exists(os.path.join(ref, "metapart")):
        return None
    buf = mp.synth(2, n_reads=sample_reads)
    d = tempfile.mkdtemp(prefix="mp_cpu_")
    try:
        fq = os.path.join(d, "in.fastq")
        with open(fq, "wb") as f:
            f.write(memoryview(buf.view()))
        buf.free()
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        C = 4 * threads
        walls, lines = [], []
        for _ in range(runs):
            od = os.path.join(d, "out")
            shutil.rmtree(od, ignore_errors=True)
            os.makedirs(os.path.join(od, "Rank0"))
            t0 = time.time()
            line = ""
            for cmd in ([os.path.join(ref, "genFastqIdx"), str(C), "0", "1", "0", fq],
                        [os.path.join(ref, "genHistograms_metapart"), str(K), "0", "1", "0", fq],
                        [os.path.join(ref, "metapart"), "-o", od, "2000000000", str(K), "1", str(threads),
                         "0", "1", "0", fq]):
                r = subprocess.run(cmd, cwd=d, env=env, stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True,
                                   timeout=600)
                if r.returncode != 0:
                    return None
                for ln in r.stdout.splitlines():
                    if "FQRead" in ln:
                        line = ln.strip()
            walls.append(time.time() - t0)
            lines.append(line)
        med = statistics.median(walls)
        return {"value": sample_reads / med, "unit": "reads/s", "cores": threads, "kind": "reference",
                "cpu_model": cpu_model(), "runs_s": walls,
                "sample": "first %d reads of the same synthetic workload (k=%d, %d chunks, S=1, T=%d, OMP_NUM_THREADS=%d), "
                          "genFastqIdx+genHistograms_metapart+metapart wall, median of %d runs: %.2f s"
                          % (sample_reads, K, C, threads, threads, runs, med),
                "stage_line": lines[walls.index(med)] if med in walls else lines[0]}
    finally:
        shutil.rmtree(d, ignore_errors=True)


def tools_wall(buf, reads: int, K: int, chunks: int):
    """The drop-in tools end to end, as a user runs them: genFastqIdx,
    genHistograms_metapart and metapart (S=1, T=1) on the workload's FASTQ file
    (written first, so it is in the page cache), partitioned files written to
    local disk.  Each tool is a process of its own (GPU runtime start-up
    included); the wall time of each and metapart's stage line."""
    bindir = os.path.join(ROOT, "metapart_amd", "bin")
    d = tempfile.mkdtemp(prefix="mp_tools_")
    try:
        fq = os.path.join(d, "in.fastq")
        with open(fq, "wb") as f:
            f.write(memoryview(buf.view()))
        os.makedirs(os.path.join(d, "out", "Rank0"))
        walls, line, reads_line = [], "", ""
        for cmd in ([os.path.join(bindir, "genFastqIdx"), str(chunks), "0", "1", "0", "in.fastq"],
                    [os.path.join(bindir, "genHistograms_metapart"), str(K), "0", "1", "0", "in.fastq"],
                    [os.path.join(bindir, "metapart"), "-o", "./out", "2000000000", str(K), "1", "1", "0", "1", "0",
                     "in.fastq"]):
            t0 = time.time()
            r = subprocess.run(cmd, cwd=d, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=900)
            walls.append(time.time() - t0)
            if r.returncode != 0:
                return {"error": "%s failed: %s" % (os.path.basename(cmd[0]), r.stdout[-500:])}
            for ln in r.stdout.splitlines():
                if "FQRead" in ln:
                    line = ln.strip()
                if "reads/s" in ln:
                    reads_line = ln.strip()
        out_bytes = sum(os.path.getsize(os.path.join(d, "out", "Rank0", f))
                        for f in os.listdir(os.path.join(d, "out", "Rank0")))
        total = sum(walls)
        return {"reads_per_s": reads / total, "wall_s": total,
                "tool_wall_s": {"genFastqIdx": walls[0], "genHistograms_metapart": walls[1], "metapart": walls[2]},
                "metapart_stage_line": line, "metapart_summary": reads_line, "out_bytes": out_bytes,
                "setup": "FASTQ in the page cache, %d chunks, S=1, T=1, output files on local disk" % chunks}
    finally:
        shutil.rmtree(d, ignore_errors=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--reads", type=int, default=20_000_000)
    ap.add_argument("--k", type=int, default=31)
    ap.add_argument("--chunks", type=int, default=16)
    ap.add_argument("--cpu-sample", type=int, default=2_000_000)
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: the cores this process may use, at most 16")
    ap.add_argument("--cpu-runs", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-tools", action="store_true", help="skip the end-to-end run of the drop-in tools")
    ap.add_argument("--no-min3", action="store_true", help="skip the 3-minimizer variant")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched 64-sample mix")
    ap.add_argument("--batch-samples", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=1)
    ap.add_argument("--no-graph", action="store_true", help="launch the batch's radix passes one by one")
    ap.add_argument("--no-serve", action="store_true", help="skip the served workload")
    ap.add_argument("--serve-jobs", type=int, default=200)
    ap.add_argument("--serve-engines", type=int, default=3)
    ap.add_argument("--serve-load", type=float, default=0.7,
                    help="arrival rate as a fraction of the plain path's job rate at the mean job")
    ap.add_argument("--no-c3", action="store_true", help="skip the large out-of-core share (100M reads)")
    ap.add_argument("--c3-reads", type=int, default=100_000_000)
    ap.add_argument("--copy-probe", action="store_true",
                    help="run one 1 GiB device copy first (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import metapart_amd as mp

    # CPU baseline first, before this process touches the GPU
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(min(args.cpu_sample, args.reads), args.k, args.cpu_threads or cpu_threads(),
                           args.cpu_runs)
        print("cpu baseline done", file=sys.stderr, flush=True)

    t0 = time.time()
    buf = mp.synth(2, n_reads=args.reads)
    gen_s = time.time() - t0
    host = buf.view()
    tools = None if args.no_tools else tools_wall(buf, args.reads, args.k, args.chunks)
    print("tools done", file=sys.stderr, flush=True)

    import torch
    dev = torch.device("cuda:0")
    d_fq = torch.empty(buf.nbytes + 64, dtype=torch.uint8, device=dev)
    d_fq[: buf.nbytes].copy_(torch.from_numpy(host))
    torch.cuda.synchronize()
    if args.copy_probe:
        src = torch.empty(1 << 30, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    def step(ctx, min3=False):
        t = {}
        a = time.perf_counter()
        cofs = mp.index_buffer(buf, args.chunks)
        t["index"] = time.perf_counter() - a
        a = time.perf_counter()
        ctx.load_device(d_fq.data_ptr(), buf.nbytes, cofs)
        t["load"] = time.perf_counter() - a
        a = time.perf_counter()
        r = mp.partition(ctx, S=1, T=1, out_prefix=None)
        t["partition"] = time.perf_counter() - a
        return r, t

    results = {}
    for variant in (["plain"] if args.no_min3 else ["plain", "min3"]):
        ctx = mp.Context(k=args.k, min3=(variant == "min3"))
        for _ in range(args.warmup):
            step(ctx)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        last = None
        for _ in range(args.steps):
            last = step(ctx)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        r, tstage = last
        sc = r.scans[0]
        results[variant] = {"ms_per_step": 1e3 * dt / args.steps, "reads_per_s": args.reads * args.steps / dt,
                            "scan": sc, "finish": r.finish, "host_stage_s": tstage, "out_bytes": r.out_bytes}
        ctx.close()
    print("timed steps done", file=sys.stderr, flush=True)

    plain = results["plain"]
    sc = plain["scan"]
    n = int(sc["tuples"])
    N = args.reads
    passes = int(sc["sort_passes"])
    pass_ms = float(sc["ms_sort_pass_avg"])
    per_pass_bytes = 2 * TUPLE_BYTES * n  # read and write 13 B per tuple
    achieved = per_pass_bytes / (pass_ms * 1e-3) / 1e9 if pass_ms > 0 else 0.0
    # the local pass (when it ran) reads every key and moves only the tuples not
    # yet at their rank: their label and adj are read and the tuple written
    local = float(sc["ms_sort_local"]) > 0
    moved = int(sc["sort_local_moved"])
    local_bytes = n * 8 + moved * (5 + TUPLE_BYTES) if local else 0
    full = passes - (1 if local else 0)  # radix passes over the whole array
    gen_ms = float(sc["ms_sort_gen"])
    bases = N * 150
    gen_bytes = n * TUPLE_BYTES + bases * 3 // 8 + bases  # tuples written; 2-bit + 'N' streams and sequence bytes read
    if gen_ms > 0:  # the first pass generates the tuples: it writes them and reads the reads
        sort_bytes = n * TUPLE_BYTES + n * 2 * (full - 1) * TUPLE_BYTES + local_bytes
    else:  # enumerated tuples: one histogram read (k-mers), then every pass reads and writes them
        sort_bytes = n * (8 + 2 * full * TUPLE_BYTES) + local_bytes
    sort_ms = float(sc["ms_sort"])
    traffic = measured_traffic(args.reads)

    def stage(ms, nbytes, what):
        gbs = nbytes / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
        return {"ms": ms, "algorithmic_bytes": int(nbytes), "GB_per_s": gbs, "frac_of_peak": gbs / HBM_PEAK_GBS,
                "bytes": what}

    stages = {
        "sort_generating_pass": stage(gen_ms, gen_bytes, "13 B per tuple written; 0.375 B per base of 2-bit and 'N' "
                                                          "words and 1 B per base of FASTQ sequence read"),
        "sort_scatter_pass": stage(pass_ms, per_pass_bytes, "13 B per tuple read and written (one pass, the kernel "
                                                            "alone)"),
        "sort_digit_counts": stage(float(sc["ms_sort_count"]), 0, "all counted passes: per-tile digit counts and "
                                                                  "offsets (no algorithmic bytes)"),
        "sort_local_pass": stage(float(sc["ms_sort_local"]), local_bytes, "8 B key per tuple read; 5 B read and 13 B "
                                                                           "written per tuple moved"),
        "filter": stage(float(sc["ms_filter"]), 9 * n, "8 B key and 1 B adj per tuple read"),
        "union": stage(float(sc["ms_union"]), n // 4 + 4 * n, "run bitmaps and the labels (random parent "
                                                              "accesses not counted)"),
    }
    out = {
        "metric": "reads partitioned/s at k=31 on 1 GPU (plain path); % HBM peak in sort",
        "value": plain["reads_per_s"],
        "unit": "reads/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": plain["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded generator, FASTQ resident in HBM)",
        "config": {"workload": "200 genomes 1-5 Mb GC U(0.3,0.7) + 10x5 kb repeats; %d SE reads x 150 bp, 1%% error; "
                               "k=%d m=10, 16 chunks, S=1, T=1, plain" % (args.reads, args.k),
                   "reads": args.reads, "k": args.k, "chunks": args.chunks, "scans": 1, "segments": 1},
        "roofline": {"bound": "hbm", "kernel": "onesweep_kernel (one radix scatter pass)", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic["bytes"] if traffic else None,
                     "traffic_source": traffic["source"] if traffic else None,
                     "bytes_per_launch": per_pass_bytes, "launch_ms": pass_ms, "tuples": n},
        "sort": {"passes": passes, "digit_bits": 8, "tuple_bytes": TUPLE_BYTES, "algorithmic_bytes": sort_bytes,
                 "ms": sort_ms, "hbm_frac": sort_bytes / (sort_ms * 1e-3) / 1e12 / (HBM_PEAK_GBS / 1e3)
                 if sort_ms > 0 else 0.0},
        "stages": stages,
        "stages_ms": {k: sc[k] for k in ("ms_enumerate", "ms_sort", "ms_sort_gen", "ms_sort_local", "ms_sort_count",
                                         "ms_filter", "ms_union")},
        "finish": plain["finish"],
        "host_stage_s": plain["host_stage_s"],
        "tools_wall": tools,
        "cpu_baseline": cpu,
        "generate_s": gen_s,
    }
    if "min3" in results:
        m3 = results["min3"]
        out["variants"] = {"min3": {"reads_per_s": m3["reads_per_s"], "ms_per_step": m3["ms_per_step"],
                                    "tuples": int(m3["scan"]["tuples"]), "finish": m3["finish"],
                                    "host_stage_s": m3["host_stage_s"],
                                    "stages_ms": {k: m3["scan"][k] for k in ("ms_enumerate", "ms_sort", "ms_filter",
                                                                             "ms_union")}}}
    del d_fq
    buf.free()
    torch.cuda.empty_cache()
    if not args.no_batch:
        out["batched"] = batch_section(args, mp, torch, dev)
        torch.cuda.empty_cache()
        print("batched done", file=sys.stderr, flush=True)
    if not args.no_serve:
        out["served"] = served_section(args, mp)
        print("served done", file=sys.stderr, flush=True)
    if not args.no_c3:
        out["out_of_core"] = c3_section(args, mp, torch)
        print("out-of-core done", file=sys.stderr, flush=True)
    print(json.dumps(out))


def c3_section(args, mp, torch):
    """1000 genomes of 0.5-8 Mb (log-normal sigma 1.5), 150 bp reads, 1 % errors,
    k = 31: the FASTQ comes from host memory through pinned staging, and the
    number of m-mer bucket passes is the fewest whose tuples fit the HBM budget."""
    from metapart_amd import tune
    t0 = time.time()
    buf = mp.synth(3, n_reads=args.c3_reads)
    gen_s = time.time() - t0
    cofs = mp.index_buffer(buf, 64)
    free, _ = torch.cuda.mem_get_info()
    budget = int(free * 0.92)
    ctx = mp.Context(k=args.k, hbm_budget=budget)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.load(buf, cofs)
    t_load = time.perf_counter() - t0
    t1 = time.perf_counter()
    hist, _, kmers, _ = ctx.count()
    bases = kmers + ctx.N * (args.k - 1)
    S, need = tune.scans_for_budget(hist, ctx.N, buf.nbytes, bases, budget)
    r = mp.partition(ctx, S=S, T=1, out_prefix=None, hist=hist)
    t_part = time.perf_counter() - t1
    out = {"workload": "%d SE reads x 150 bp, 1000 genomes 0.5-8 Mb, log-normal(1.5), 1%% error, k=%d, 64 chunks, "
                       "T=1, scans from the HBM budget" % (ctx.N, args.k),
           "reads": ctx.N, "fastq_bytes": buf.nbytes, "kmers": kmers, "scans": S,
           "scan_tuples": [int(x["tuples"]) for x in r.scans], "scan_parts": [int(x["parts"]) for x in r.scans],
           "budget_bytes": budget, "planned_bytes": need,
           "device_bytes": ctx.device_bytes, "ms_load_from_host": 1e3 * t_load, "ms_partition": 1e3 * t_part,
           "reads_per_s": ctx.N / t_part, "reads_per_s_with_upload": ctx.N / (t_part + t_load),
           "stages_ms": [{k: x[k] for k in ("ms_enumerate", "ms_sort", "ms_filter", "ms_union")} for x in r.scans],
           "finish": r.finish, "generate_s": gen_s}
    ctx.close()
    buf.free()
    return out


def _pct(xs, q):
    import numpy as np
    return float(np.percentile(np.asarray(xs), q)) if xs else 0.0


def served_section(args, mp):
    """Poisson-arriving partition jobs (job j: seed 5000 + j, its own pool of
    20-100 genomes of 0.5-3 Mb, log-uniform 50k-2M reads of 150 bp, 1 % errors)
    from host memory, at a rate of serve_load x the plain path's job rate at
    the mean job (529 k reads): the request loop (engines, HBM-budget
    admission, continuous batching) against the plain per-job path (one
    context, jobs one at a time in arrival order), same arrivals; then every
    job queued at once (saturation) against the plain path back to back."""
    import numpy as np
    params = mp.job_params(n_jobs=args.serve_jobs, seed=7)
    t0 = time.time()
    bufs = mp.synth_many(params, threads=8)
    gen_s = time.time() - t0
    reads = [p["n_reads"] for p in params]
    warm = mp.synth(0, n_reads=20000, seed=5, read_len=150)
    # the plain path's job rate at the mean job
    mean_job = mp.synth(0, seed=4999, n_genomes=60, genome_len_min=500_000, genome_len_max=3_000_000, abund_sigma=1.0,
                        n_reads=529_000, read_len=150, err_rate=0.01)
    ctx = mp.Context(k=args.k)
    ctx.load(warm, [0, warm.nbytes])
    mp.partition(ctx, S=1, T=1, build_output=False)
    ts = []
    for _ in range(4):
        a = time.perf_counter()
        ctx.load(mean_job, [0, mean_job.nbytes])
        mp.partition(ctx, S=1, T=1, build_output=False)
        ts.append(time.perf_counter() - a)
    mean_job.free()
    capacity = 1.0 / float(np.median(ts))
    rate = args.serve_load * capacity
    rng = np.random.default_rng(11)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=len(bufs)))

    # plain per-job path
    lat_plain = []
    start = time.perf_counter()
    for i, b in enumerate(bufs):
        now = time.perf_counter() - start
        if now < arrivals[i]:
            time.sleep(arrivals[i] - now)
        ctx.load(b, [0, b.nbytes])
        mp.partition(ctx, S=1, T=1, build_output=False)
        lat_plain.append(time.perf_counter() - start - arrivals[i])
    plain_span = time.perf_counter() - start
    ctx.close()

    # served
    E = args.serve_engines
    srv = mp.Server(k=args.k, engines=E)
    srv.wait(srv.submit(warm))
    jobs = []
    start = time.perf_counter()
    for i, b in enumerate(bufs):
        now = time.perf_counter() - start
        if now < arrivals[i]:
            time.sleep(arrivals[i] - now)
        jobs.append(srv.submit(b))
    res = [srv.wait(j)[0] for j in jobs]
    lat = [r["t_done"] - r["t_submit"] for r in res]
    span = max(r["t_done"] for r in res) - min(r["t_submit"] for r in res)
    # saturation: every job queued at once, against the plain path's back-to-back rate
    t0 = time.perf_counter()
    sat = [srv.submit(b) for b in bufs]
    sres = [srv.wait(j)[0] for j in sat]
    sat_s = time.perf_counter() - t0
    srv.close()
    ctx = mp.Context(k=args.k)
    t0 = time.perf_counter()
    for b in bufs:
        ctx.load(b, [0, b.nbytes])
        mp.partition(ctx, S=1, T=1, build_output=False)
    plain_sat_s = time.perf_counter() - t0
    ctx.close()
    for b in bufs:
        b.free()
    warm.free()
    return {"workload": "%d jobs, log-uniform 50k-2M SE reads of 150 bp, 1%% error, k=%d, job j: fresh pool of 20-100 "
                        "genomes of 0.5-3 Mb seeded 5000+j; FASTQ in host memory; Poisson arrivals at %.1f jobs/s = "
                        "%.2f x the plain path's %.1f jobs/s at a 529k-read job"
                        % (len(bufs), args.k, rate, args.serve_load, capacity),
            "jobs": len(bufs), "reads": int(sum(reads)), "rate_jobs_per_s": rate,
            "plain_capacity_jobs_per_s": capacity, "mean_job_ms": 1e3 / capacity,
            "served": {"p50_ms": 1e3 * _pct(lat, 50), "p99_ms": 1e3 * _pct(lat, 99),
                       "reads_per_s": sum(reads) / span, "engines": E,
                       "mean_batch_jobs": float(np.mean([r["batch_jobs"] for r in res]))},
            "plain": {"p50_ms": 1e3 * _pct(lat_plain, 50), "p99_ms": 1e3 * _pct(lat_plain, 99),
                      "reads_per_s": sum(reads) / plain_span},
            "saturation": {"served_reads_per_s": sum(reads) / sat_s, "plain_reads_per_s": sum(reads) / plain_sat_s,
                           "speedup": plain_sat_s / sat_s,
                           "served_mean_batch_jobs": float(np.mean([r["batch_jobs"] for r in sres])),
                           "setup": "all %d jobs queued at once (served) / run back to back (plain)" % len(bufs)},
            "generate_s": gen_s}


def batch_section(args, mp, torch, dev):
    """Plain per-sample path vs one batch over the same HBM-resident mix."""
    import numpy as np
    params = mp.mix_params(n_samples=args.batch_samples, seed=4)
    t0 = time.time()
    bufs = mp.synth_many(params, threads=8)
    gen_s = time.time() - t0
    sizes = [b.nbytes for b in bufs]
    ofs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    total = int(ofs[-1])
    d = torch.empty(total + 64, dtype=torch.uint8, device=dev)
    # the plain path takes each sample as its own 16-byte aligned buffer
    aofs = np.concatenate([[0], np.cumsum([(n + 255) // 256 * 256 for n in sizes])]).astype(np.uint64)
    da = torch.empty(int(aofs[-1]) + 64, dtype=torch.uint8, device=dev)
    for b, o, a in zip(bufs, ofs[:-1], aofs[:-1]):
        h = torch.from_numpy(b.view())
        d[int(o): int(o) + b.nbytes].copy_(h)
        da[int(a): int(a) + b.nbytes].copy_(h)
    torch.cuda.synchronize()
    for b in bufs:
        b.free()
    reads = sum(p["n_reads"] for p in params)
    base = d.data_ptr()

    ctx = mp.Context(k=args.k)

    def plain_step():
        Ws = []
        for i in range(len(sizes)):
            ctx.load_device(da.data_ptr() + int(aofs[i]), sizes[i], [0, sizes[i]])
            Ws.append(mp.partition(ctx, S=1, T=1, out_prefix=None).W)
        return Ws

    plain_step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.batch_steps):
        Wp = plain_step()
    torch.cuda.synchronize()
    plain_s = (time.perf_counter() - t0) / args.batch_steps
    ctx.close()
    del da
    torch.cuda.empty_cache()

    # the batch's pass chain captured in a HIP graph by the warm-up step, replayed by the timed ones
    bctx = mp.Context(k=args.k, sort_graph=not args.no_graph)

    def batch_step():
        bctx.batch_load_device(base, total, ofs)
        res, st, fs = bctx.batch_partition()
        nb = bctx.batch_write(None)
        return res, st, fs, nb

    batch_step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.batch_steps):
        res, st, fs, nb = batch_step()
    torch.cuda.synchronize()
    batch_s = (time.perf_counter() - t0) / args.batch_steps
    bctx.close()
    equal = all(np.array_equal(w, r["W"]) for w, r in zip(Wp, res))
    return {"workload": "%d samples, sample i: own pool of 20-300 genomes of 0.1-2 Mb (GC U(0.3,0.7)) seeded 100+i, "
                        "log-uniform 10k-5M SE reads of 100-150 bp, 0.5-2%% error, k=%d, one chunk/scan/segment per "
                        "sample" % (len(sizes), args.k),
            "samples": len(sizes), "reads": reads, "tuples": int(st.tuples),
            "plain_reads_per_s": reads / plain_s, "batched_reads_per_s": reads / batch_s,
            "speedup": plain_s / batch_s, "ms_plain": 1e3 * plain_s, "ms_batched": 1e3 * batch_s,
            "plans_equal": bool(equal), "batch_stages_ms": {k: getattr(st, k) for k in
                                                            ("ms_enumerate", "ms_sort", "ms_filter", "ms_union")},
            "batch_sort_passes": st.sort_passes, "batch_sort_pass_ms": st.ms_sort_pass_avg,
            "sort_graph": {0: "launched pass by pass", 1: "captured", 2: "replayed"}.get(st.sort_graph),
            "generate_s": gen_s}


if __name__ == "__main__":
    main()
