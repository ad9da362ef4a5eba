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

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--reads N] [--no-cpu-baseline]
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
    ap.add_argument("--min3", action="store_true", help="also time the 3-minimizer variant")
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
    for variant in (["plain", "min3"] if args.min3 else ["plain"]):
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
    sort_bytes = n * (8 + 2 * passes * TUPLE_BYTES)
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
        "stages_ms": {k: sc[k] for k in ("ms_enumerate", "ms_sort", "ms_filter", "ms_union")},
        "finish": plain["finish"],
        "host_stage_s": plain["host_stage_s"],
        "cpu_baseline": cpu,
        "generate_s": gen_s,
    }
    if "min3" in results:
        out["variants"] = {"min3": {"reads_per_s": results["min3"]["reads_per_s"],
                                    "ms_per_step": results["min3"]["ms_per_step"]}}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
