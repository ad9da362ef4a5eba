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

The same line also reports the 3-minimizer variant on the same reads
("variants") and the batched form ("batched"): a 64-sample mix (seeded genome
pool per sample of 20-300 genomes of 0.1-1 Mb, log-uniform 10k-5M reads of
100-150 bp, 0.5-2 % errors, k = 31) resident in HBM, run once through the
plain per-sample path (one context, samples one after another, each a full
histogram/plan/scan/finish/writer run) and once as one batch; the plans of
every sample must agree between the two.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--reads N] [--no-cpu-baseline]
                       [--no-min3] [--no-batch] [--batch-samples 64]
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


def cpu_baseline(sample_reads: int, K: int, cores: int = 1):
    """The reference implementation (built from its sources into oracle/_ref) on
    the first sample_reads reads of the workload, one thread: genFastqIdx +
    genHistograms_metapart + metapart wall time."""
    import metapart_amd as mp
    ref = os.path.join(ROOT, "oracle", "_ref")
    if not os.path.exists(os.path.join(ref, "metapart")):
        return None
    buf = mp.synth(2, n_reads=sample_reads)
    d = tempfile.mkdtemp(prefix="mp_cpu_")
    try:
        fq = os.path.join(d, "in.fastq")
        with open(fq, "wb") as f:
            f.write(buf.tobytes())
        buf.free()
        os.makedirs(os.path.join(d, "out", "Rank0"))
        env = dict(os.environ, OMP_NUM_THREADS=str(cores))
        t0 = time.time()
        for cmd in ([os.path.join(ref, "genFastqIdx"), "16", "0", "1", "0", fq],
                    [os.path.join(ref, "genHistograms_metapart"), str(K), "0", "1", "0", fq],
                    [os.path.join(ref, "metapart"), "-o", os.path.join(d, "out"), "2000000000", str(K), "1", str(cores),
                     "0", "1", "0", fq]):
            r = subprocess.run(cmd, cwd=d, env=env, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, timeout=600)
            if r.returncode != 0:
                return None
        dt = time.time() - t0
        return {"value": sample_reads / dt, "unit": "reads/s", "cores": cores, "kind": "reference",
                "sample": "first %d reads of the same synthetic workload (k=%d, 16 chunks, S=1, T=%d), "
                          "genFastqIdx+genHistograms_metapart+metapart wall %.2f s" % (sample_reads, K, cores, dt)}
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
    ap.add_argument("--cpu-sample", type=int, default=600_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-min3", action="store_true", help="skip the 3-minimizer variant")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched 64-sample mix")
    ap.add_argument("--batch-samples", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=1)
    ap.add_argument("--no-serve", action="store_true", help="skip the served workload")
    ap.add_argument("--c3", action="store_true",
                    help="also run the large out-of-core share (100M reads; needs ~45 GB of host memory)")
    ap.add_argument("--c3-reads", type=int, default=100_000_000)
    ap.add_argument("--serve-jobs", type=int, default=64)
    ap.add_argument("--serve-rate", type=float, default=60.0, help="Poisson arrival rate, jobs/s")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run one 1 GiB device copy first (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import metapart_amd as mp

    # CPU baseline first, before this process touches the GPU
    cpu = None if args.no_cpu_baseline else cpu_baseline(min(args.cpu_sample, args.reads), args.k)

    t0 = time.time()
    buf = mp.synth(2, n_reads=args.reads)
    gen_s = time.time() - t0
    host = buf.view()

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

    plain = results["plain"]
    sc = plain["scan"]
    n = int(sc["tuples"])
    passes = int(sc["sort_passes"])
    pass_ms = float(sc["ms_sort_pass_avg"])
    per_pass_bytes = 2 * TUPLE_BYTES * n  # read and write 13 B per tuple
    achieved = per_pass_bytes / (pass_ms * 1e-3) / 1e9 if pass_ms > 0 else 0.0
    # the local pass (when it ran) reads every key and moves only the tuples not
    # yet at their rank: their label and adj are read and the tuple written
    local = float(sc["ms_sort_local"]) > 0
    local_bytes = n * 8 + int(sc["sort_local_moved"]) * (5 + TUPLE_BYTES) if local else 0
    full = passes - (1 if local else 0)  # radix passes over the whole array
    if float(sc["ms_sort_gen"]) > 0:  # the first pass generates the tuples: it writes them and reads the reads
        sort_bytes = n * TUPLE_BYTES + n * 2 * (full - 1) * TUPLE_BYTES + local_bytes
    else:  # enumerated tuples: one histogram read (k-mers), then every pass reads and writes them
        sort_bytes = n * (8 + 2 * full * TUPLE_BYTES) + local_bytes
    sort_ms = float(sc["ms_sort"])
    traffic = measured_traffic(args.reads)
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
        "stages_ms": {k: sc[k] for k in ("ms_enumerate", "ms_sort", "ms_sort_gen", "ms_sort_local", "ms_filter",
                                         "ms_union")},
        "finish": plain["finish"],
        "host_stage_s": plain["host_stage_s"],
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
    if not args.no_batch:
        del d_fq
        buf.free()
        torch.cuda.empty_cache()
        out["batched"] = batch_section(args, mp, torch, dev)
    if not args.no_serve:
        out["served"] = served_section(args, mp)
    if args.c3:
        out["out_of_core"] = c3_section(args, mp, torch)
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
    """Poisson-arriving jobs from host memory: the request loop (2 engines,
    continuous batching under an HBM budget) against the plain per-job path
    (one context, jobs one at a time in arrival order), same arrivals."""
    import numpy as np
    params = mp.job_params(n_jobs=args.serve_jobs, seed=7)
    t0 = time.time()
    bufs = [mp.synth(0, **p) for p in params]
    gen_s = time.time() - t0
    reads = [p["n_reads"] for p in params]
    rng = np.random.default_rng(11)
    arrivals = np.cumsum(rng.exponential(1.0 / args.serve_rate, size=len(bufs)))
    warm = mp.synth(0, n_reads=20000, seed=5, read_len=150)

    # plain per-job path
    ctx = mp.Context(k=args.k)
    ctx.load(warm, [0, warm.nbytes])
    mp.partition(ctx, S=1, T=1, build_output=False)
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
    srv = mp.Server(k=args.k, engines=2)
    srv.wait(srv.submit(warm))
    jobs = []
    start = time.perf_counter()
    for i, b in enumerate(bufs):
        now = time.perf_counter() - start
        if now < arrivals[i]:
            time.sleep(arrivals[i] - now)
        jobs.append(srv.submit(b))
    res = [srv.wait(j)[0] for j in jobs]
    srv.close()
    lat = [r["t_done"] - r["t_submit"] for r in res]
    span = max(r["t_done"] for r in res) - min(r["t_submit"] for r in res)
    for b in bufs:
        b.free()
    return {"workload": "%d jobs, log-uniform 50k-2M SE reads of 150 bp, 1%% error, k=%d, fresh genome pool per job "
                        "(20-100 genomes of 0.1-1 Mb), FASTQ in host memory, Poisson arrivals at %.0f jobs/s"
                        % (len(bufs), args.k, args.serve_rate),
            "jobs": len(bufs), "reads": int(sum(reads)), "rate_jobs_per_s": args.serve_rate,
            "served": {"p50_ms": 1e3 * _pct(lat, 50), "p99_ms": 1e3 * _pct(lat, 99),
                       "reads_per_s": sum(reads) / span, "engines": 2,
                       "mean_batch_jobs": float(np.mean([r["batch_jobs"] for r in res]))},
            "plain": {"p50_ms": 1e3 * _pct(lat_plain, 50), "p99_ms": 1e3 * _pct(lat_plain, 99),
                      "reads_per_s": sum(reads) / plain_span},
            "generate_s": gen_s}


def batch_section(args, mp, torch, dev):
    """Plain per-sample path vs one batch over the same HBM-resident mix."""
    import numpy as np
    params = mp.mix_params(n_samples=args.batch_samples, seed=4)
    t0 = time.time()
    bufs = [mp.synth(0, **p) for p in params]
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

    bctx = mp.Context(k=args.k)

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
    return {"workload": "%d samples, own genome pools (20-300 genomes of 0.1-1 Mb), log-uniform 10k-5M SE reads of "
                        "100-150 bp, 0.5-2%% error, k=%d, one chunk/scan/segment per sample" % (len(sizes), args.k),
            "samples": len(sizes), "reads": reads, "tuples": int(st.tuples),
            "plain_reads_per_s": reads / plain_s, "batched_reads_per_s": reads / batch_s,
            "speedup": plain_s / batch_s, "ms_plain": 1e3 * plain_s, "ms_batched": 1e3 * batch_s,
            "plans_equal": bool(equal), "batch_stages_ms": {k: getattr(st, k) for k in
                                                            ("ms_enumerate", "ms_sort", "ms_filter", "ms_union")},
            "batch_sort_passes": st.sort_passes, "batch_sort_pass_ms": st.ms_sort_pass_avg,
            "generate_s": gen_s}


if __name__ == "__main__":
    main()
