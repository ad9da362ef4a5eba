"""The benchmark workloads at their real sizes.

* 20 M reads of 150 bp (200 genomes of 1-5 Mb, GC U(0.3, 0.7), planted 5 kb
  repeats, 1 % errors), k = 31, 16 chunks: the drop-in tools against the
  reference's output hashes (two scans with one and two segments, and the
  3-minimizer variant with one scan), and the one-scan, one-segment run the
  benchmark times, plain and 3-minimizer: invariants and run-to-run identity.
* The first 2 M reads of the 100 M-read workload (1000 genomes of 0.5-8 Mb,
  log-normal(1.5)), 64 chunks, four scans: the tools against the reference.
* The 100 M-read workload itself, 64 chunks, four scans, one segment, with
  the FASTQ uploaded through pinned staging: invariants and identity.
* The 64-sample mix (sample seeds 100-163, 20-300 genomes of 0.1-2 Mb,
  log-uniform 10 k-5 M reads of 100-150 bp, 0.5-2 % errors) as one batch:
  every sample equals its plain run alone, and every sample of at most 200 k
  reads equals the oracle (labels and output files).
* Jobs of the served workload (job seeds 5000 + j, 20-100 genomes of
  0.5-3 Mb, log-uniform 50 k-2 M reads of 150 bp, 1 % errors) submitted over
  time to the request loop: every job equals its plain run alone, and every
  job of at most 100 k reads equals the oracle.
"""
import hashlib
import os
import time

import numpy as np
import pytest

from _helpers import diff_hashes, golden_large, make_large_input, oracle_samples, run_chain, sha

import metapart_amd as mp

LARGE = golden_large()


@pytest.fixture(scope="module")
def large_inputs(tmp_path_factory):
    d = tmp_path_factory.mktemp("large_inputs")
    made = {}

    def get(name):
        if name not in made:
            p = str(d / (name + ".fastq"))
            t0 = time.time()
            make_large_input(name, p)
            print("  generated %s (%.1f GB) in %.0f s" % (name, os.path.getsize(p) / 1e9, time.time() - t0),
                  flush=True)
            made[name] = p
        return made[name]
    return get


@pytest.mark.gpu
@pytest.mark.parametrize("case", LARGE["cases"], ids=[c["name"] for c in LARGE["cases"]])
def test_workload_tools_match_reference(case, large_inputs, tmp_path):
    print("gpu workload golden case", case["name"], flush=True)
    src = large_inputs(case["input"])
    got = run_chain("gpu", case, str(tmp_path), timeout=900, input_path=src)
    assert diff_hashes(got, case["sha256"]) == []


def _invariants(N, r):
    par = r.parent.astype(np.int64)
    assert np.all(par >= np.arange(N))  # roots are component maxima
    assert np.all(par[par] == par)      # every label is a root
    W = r.W.astype(np.int64)
    assert W[0] == 0 and W[16] == N and np.all(np.diff(W) >= 0)
    roots, sizes = np.unique(par, return_counts=True)
    assert r.finish["components"] == roots.size
    assert r.finish["largest"] == sizes.max()


@pytest.mark.gpu
@pytest.mark.parametrize("min3", [False, True], ids=["plain", "min3"])
def test_medium_workload_one_scan_properties_and_identity(min3, large_inputs):
    print("gpu 20M-read workload, S=1 T=1, %s" % ("3-minimizer" if min3 else "plain"), flush=True)
    buf = np.fromfile(large_inputs("c2"), dtype=np.uint8)
    N = 20_000_000
    digests = set()
    for rep in range(2):
        with mp.Context(k=31, min3=min3) as ctx:
            ctx.load(buf, mp.index_buffer(buf, 16))
            r = mp.partition(ctx, S=1, T=1, want_parent=True)
            assert ctx.N == N
        sc = r.scans[0]
        if not min3:
            assert sc["tuples"] == int(r.hist[-1]) == N * 120  # 150 bp, k = 31, no 'N'
        else:
            assert sc["tuples"] == int(r.hist[-1])
        _invariants(N, r)
        assert r.out_bytes == buf.nbytes  # single-end: every record once
        digests.add((hashlib.sha256(r.parent.tobytes()).hexdigest(), tuple(int(x) for x in r.W)))
        print("  rep %d: %d components, largest %d" % (rep, r.finish["components"], r.finish["largest"]), flush=True)
    assert len(digests) == 1


@pytest.mark.gpu
def test_large_workload_four_scans_properties_and_identity():
    print("gpu 100M-read workload, C=64 S=4 T=1, host upload", flush=True)
    t0 = time.time()
    buf = mp.synth(3)
    print("  generated %.1f GB in %.0f s" % (buf.nbytes / 1e9, time.time() - t0), flush=True)
    N = 100_000_000
    cofs = mp.index_buffer(buf, 64)
    digests = set()
    with mp.Context(k=31) as ctx:
        for rep in range(2):
            ctx.load(buf, cofs)
            r = mp.partition(ctx, S=4, T=1, want_parent=True)
            assert ctx.N == N
            assert sum(int(s["tuples"]) for s in r.scans) == int(r.hist[-1]) == N * 120
            assert all(int(s["parts"]) == 0 for s in r.scans)  # every scan fitted
            _invariants(N, r)
            assert r.out_bytes == buf.nbytes
            digests.add((hashlib.sha256(r.parent.tobytes()).hexdigest(), tuple(int(x) for x in r.W)))
            print("  rep %d: scans %s, %d components" % (rep, [int(s["tuples"]) for s in r.scans],
                                                         r.finish["components"]), flush=True)
    buf.free()
    assert len(digests) == 1


def _hash_dir(d):
    return {f: sha(os.path.join(d, f)) for f in sorted(os.listdir(d))}


def _write(buf, path):
    with open(path, "wb") as f:
        f.write(memoryview(buf.view()))


@pytest.mark.gpu
def test_sample_mix_batch_equals_solo_runs_and_oracle(tmp_path):
    params = mp.mix_params(n_samples=64, seed=4)
    print("gpu 64-sample mix: %d reads" % sum(p["n_reads"] for p in params), flush=True)
    bufs = mp.synth_many(params)
    K = 31
    small = [i for i, p in enumerate(params) if p["n_reads"] <= 200_000]
    # every sample alone: labels and plan; output files of the small ones
    solo = []
    with mp.Context(k=K) as ctx:
        for i, b in enumerate(bufs):
            ctx.load(b, [0, b.nbytes])
            od = None
            if i in small:
                od = tmp_path / "solo" / str(i)
                (od / "Rank0").mkdir(parents=True)
            r = mp.partition(ctx, S=1, T=1, out_prefix=str(od) if od else None, want_parent=True)
            solo.append((r.W.copy(), r.parent.astype(np.uint32), _hash_dir(str(od / "Rank0")) if od else None))
    print("  solo runs done", flush=True)
    # the oracle on every sample of at most 200 k reads
    jobs = []
    for i in small:
        d = tmp_path / "oracle" / str(i)
        d.mkdir(parents=True)
        _write(bufs[i], str(d / "in.fastq"))
        jobs.append((str(d / "in.fastq"), 1, K, 1, 1, str(d)))
    ref = oracle_samples(jobs)
    print("  oracle on %d samples done" % len(small), flush=True)
    for i, (par, files) in zip(small, ref):
        assert np.array_equal(solo[i][1], par), i
        assert solo[i][2] == files, i
    # the whole mix as one batch: labels and plans
    with mp.Context(k=K) as ctx:
        ctx.batch_load(bufs)
        res, st, fs = ctx.batch_partition()
        print("  batch: %d tuples, %d sort passes" % (st.tuples, st.sort_passes), flush=True)
        for i in range(len(bufs)):
            assert res[i]["reads"] == params[i]["n_reads"]
            assert np.array_equal(res[i]["W"], solo[i][0]), i
            assert np.array_equal(ctx.batch_parent(i), solo[i][1]), i
        # the small samples as a batch of their own, with output files
        ctx.batch_load([bufs[i] for i in small])
        res, _, _ = ctx.batch_partition()
        ctx.batch_write(str(tmp_path / "batch"))
        for j, i in enumerate(small):
            assert np.array_equal(res[j]["W"], solo[i][0]), i
            assert np.array_equal(ctx.batch_parent(j), solo[i][1]), i
            assert _hash_dir(str(tmp_path / "batch" / ("sample%d" % j) / "Rank0")) == solo[i][2], i
    for b in bufs:
        b.free()


@pytest.mark.gpu
def test_served_jobs_equal_alone_and_oracle(tmp_path):
    params = mp.job_params(n_jobs=48, seed=7)
    print("gpu served jobs: %d jobs, %d reads" % (len(params), sum(p["n_reads"] for p in params)), flush=True)
    bufs = mp.synth_many(params)
    K = 31
    alone = []
    t0 = time.perf_counter()
    with mp.Context(k=K) as ctx:
        for b in bufs:
            ctx.load(b, [0, b.nbytes])
            r = mp.partition(ctx, S=1, T=1, want_parent=True, build_output=False)
            alone.append((r.W.copy(), r.parent.astype(np.uint32)))
    per_job = (time.perf_counter() - t0) / len(bufs)
    small = [i for i, p in enumerate(params) if p["n_reads"] <= 100_000]
    jobs = []
    for i in small:
        d = tmp_path / str(i)
        d.mkdir()
        _write(bufs[i], str(d / "in.fastq"))
        jobs.append((str(d / "in.fastq"), 1, K, 1, 1, str(d)))
    ref = oracle_samples(jobs)
    for i, (par, _) in zip(small, ref):
        assert np.array_equal(alone[i][1], par), i
    # arrivals at about twice the plain path's job rate: jobs arrive while others run
    rng = np.random.default_rng(3)
    gaps = rng.exponential(per_job / 2, size=len(bufs))
    with mp.Server(k=K, engines=2) as srv:
        ids = []
        for b, g in zip(bufs, gaps):
            time.sleep(g)
            ids.append(srv.submit(b, want_parent=True))
        got = [srv.wait(j, want_parent=True) for j in ids]
    batched = 0
    for (res, par), (W, p0) in zip(got, alone):
        assert np.array_equal(res["W"], W)
        assert np.array_equal(par, p0)
        batched += res["batch_jobs"] > 1
    print("  %d of %d jobs ran with others" % (batched, len(got)), flush=True)
    for b in bufs:
        b.free()
