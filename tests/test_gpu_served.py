"""Served form: jobs submitted over time to the request loop come back with
the result of a plain run of each sample alone, whether they ran alone or
batched with others, and the HBM budget turns a job that cannot fit into an
error for that job only; 3-minimizer jobs are served as plain ones are."""
import time

import numpy as np
import pytest

from _helpers import oracle_run

import metapart_amd as mp

K = 31


def _jobs():
    specs = [dict(n_reads=40000, n_genomes=20, genome_len_min=20000, genome_len_max=60000, read_len=150),
             dict(n_reads=4000, n_genomes=5, genome_len_min=5000, genome_len_max=9000, read_len=150),
             dict(n_reads=12000, n_genomes=40, genome_len_min=4000, genome_len_max=8000, read_len=120),
             dict(n_reads=2000, n_genomes=2, genome_len_min=3000, genome_len_max=6000, read_len=100),
             dict(n_reads=20000, n_genomes=10, genome_len_min=10000, genome_len_max=30000, read_len=150),
             dict(n_reads=6000, n_genomes=8, genome_len_min=4000, genome_len_max=8000, read_len=150)]
    return [mp.synth(0, seed=900 + i, err_rate=0.01, **sp) for i, sp in enumerate(specs)]


def _alone(buf, min3=False):
    with mp.Context(k=K, min3=min3) as ctx:
        ctx.load(buf, [0, buf.nbytes])
        r = mp.partition(ctx, S=1, T=1, want_parent=True, build_output=False)
    return r.W, r.parent.astype(np.uint32)


@pytest.mark.gpu
def test_served_results_equal_alone_and_in_the_mix():
    print("gpu served: 6 jobs, one engine, then two", flush=True)
    bufs = _jobs()
    alone = [_alone(b) for b in bufs]
    for engines in (1, 2):
        with mp.Server(k=K, engines=engines) as srv:
            jobs = [srv.submit(bufs[0], want_parent=True)]
            jobs += [srv.submit(b, want_parent=True) for b in bufs[1:]]  # arrive while job 0 runs
            got = [srv.wait(j, want_parent=True) for j in jobs]
        sizes = []
        for (res, par), (W, p0) in zip(got, alone):
            assert np.array_equal(res["W"], W)
            assert np.array_equal(par, p0)
            assert res["t_submit"] <= res["t_start"] <= res["t_done"]
            sizes.append(res["batch_jobs"])
        assert max(sizes) >= 2, sizes  # later arrivals were batched
    # every job alone through the server (batches of one)
    with mp.Server(k=K, engines=1, max_jobs_per_batch=1) as srv:
        jobs = [srv.submit(b) for b in bufs]
        for j, (W, _) in zip(jobs, alone):
            res, _ = srv.wait(j)
            assert res["batch_jobs"] == 1
            assert np.array_equal(res["W"], W)


@pytest.mark.gpu
def test_served_budget_rejects_only_the_oversized_job():
    print("gpu served: HBM budget", flush=True)
    small = mp.synth(0, n_reads=2000, seed=77)
    big = mp.synth(0, n_reads=200000, seed=78, read_len=150)
    W0, _ = _alone(small)
    with mp.Server(k=K, engines=1, hbm_budget=48 << 20) as srv:
        jb = srv.submit(big)
        js = srv.submit(small)
        with pytest.raises(mp.MPError) as e:
            srv.wait(jb)
        assert e.value.code == -2
        res, _ = srv.wait(js)
        assert np.array_equal(res["W"], W0)


@pytest.mark.gpu
def test_served_min3_jobs_equal_alone_and_oracle(tmp_path):
    """3-minimizer jobs: batched by the server, each equals its run alone, and
    two of them equal the oracle."""
    print("gpu served: 3-minimizer jobs", flush=True)
    specs = [(8000, 6), (2000, 2), (12000, 12), (4000, 3)]
    bufs = [mp.synth(0, seed=950 + i, err_rate=0.01, n_reads=n, n_genomes=g, genome_len_min=4000,
                     genome_len_max=20000, read_len=150) for i, (n, g) in enumerate(specs)]
    alone = [_alone(b, min3=True) for b in bufs]
    with mp.Server(k=K, engines=1, min3=True) as srv:
        jobs = [srv.submit(b, want_parent=True) for b in bufs]  # jobs 1-3 arrive while job 0 runs
        got = [srv.wait(j, want_parent=True) for j in jobs]
    for (res, par), (W, p0) in zip(got, alone):
        assert np.array_equal(res["W"], W)
        assert np.array_equal(par, p0)
    assert max(res["batch_jobs"] for res, _ in got) >= 2
    for i in (0, 2):
        p = str(tmp_path / ("s%d.fastq" % i))
        open(p, "wb").write(bufs[i].tobytes())
        ref = oracle_run(p, 1, K, 1, 1, min3=True, out_dir=str(tmp_path / ("o%d" % i)))
        assert np.array_equal(alone[i][1], ref["parent"].astype(np.uint32))
