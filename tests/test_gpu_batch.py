"""Batched form: a batch of independent samples partitioned in one launch
sequence gives, for every sample, the labels, plan and output files of a plain
run of that sample alone (one chunk, one scan, one segment), which equal the
oracle's on the same sample."""
import hashlib
import os

import numpy as np
import pytest

from _helpers import oracle_samples

import metapart_amd as mp

K = 31


def _samples():
    # uneven sizes, own genome pools, different read lengths and error rates
    specs = [dict(n_reads=2000, n_genomes=3, genome_len_min=4000, genome_len_max=8000, read_len=100, err_rate=0.005),
             dict(n_reads=30000, n_genomes=12, genome_len_min=5000, genome_len_max=20000, read_len=150, err_rate=0.01),
             dict(n_reads=8000, n_genomes=40, genome_len_min=4000, genome_len_max=8000, read_len=120, err_rate=0.02),
             dict(n_reads=2, n_genomes=1, genome_len_min=3000, genome_len_max=3000, read_len=100, err_rate=0.0),
             dict(n_reads=12000, n_genomes=6, genome_len_min=20000, genome_len_max=40000, read_len_min=60,
                  read_len_max=150, err_rate=0.01, n_rate=0.002),
             dict(n_reads=20000, n_genomes=40, genome_len_min=4000, genome_len_max=8000, abund_sigma=1.0,
                  read_len=100, err_rate=0.005)]
    return [mp.synth(0, seed=500 + i, **sp) for i, sp in enumerate(specs)]


def _hash_dir(d):
    out = {}
    for root, _, files in os.walk(d):
        for f in files:
            p = os.path.join(root, f)
            out[os.path.relpath(p, d)] = hashlib.sha256(open(p, "rb").read()).hexdigest()
    return out


def _solo(buf, out_dir):
    with mp.Context(k=K) as ctx:
        ctx.load(buf, [0, buf.nbytes])
        os.makedirs(os.path.join(out_dir, "Rank0"), exist_ok=True)
        r = mp.partition(ctx, S=1, T=1, out_prefix=out_dir, want_parent=True)
    return r


@pytest.fixture(scope="module")
def oracle_ref(tmp_path_factory):
    """The oracle on every sample alone: (labels, {output file: sha256})."""
    d = tmp_path_factory.mktemp("oracle")
    jobs = []
    for i, b in enumerate(_samples()):
        sd = d / str(i)
        sd.mkdir()
        (sd / "in.fastq").write_bytes(b.tobytes())
        jobs.append((str(sd / "in.fastq"), 1, K, 1, 1, str(sd)))
    return oracle_samples(jobs)


# segmented: every sample's tuples sorted in tiles of their own, from its own
# position (the default); sample_pass: look-back passes over all samples, then
# a last pass on the sample of the label (MP_SORT_LOOKBACK=1)
@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["segmented", "sample_pass"])
def test_batch_equals_solo_runs(mode, tmp_path, monkeypatch, oracle_ref):
    print("gpu batch: 6 samples vs solo runs and the oracle, %s sort" % mode, flush=True)
    if mode == "sample_pass":
        monkeypatch.setenv("MP_SORT_LOOKBACK", "1")
    bufs = _samples()
    solo = []
    for i, b in enumerate(bufs):
        solo.append(_solo(b, str(tmp_path / ("solo%d" % i))))
        par, files = oracle_ref[i]
        assert np.array_equal(solo[-1].parent.astype(np.uint32), par), i
        rank0 = tmp_path / ("solo%d" % i) / "Rank0"
        assert {f: hashlib.sha256((rank0 / f).read_bytes()).hexdigest() for f in sorted(os.listdir(rank0))} == files, i
    with mp.Context(k=K) as ctx:
        ctx.batch_load(bufs)
        res, st, fs = ctx.batch_partition()
        assert st.tuples == sum(int(s.scans[0]["tuples"]) for s in solo)
        nb = ctx.batch_write(str(tmp_path / "batch"))
        for i, (b, s) in enumerate(zip(bufs, solo)):
            assert res[i]["reads"] == len(s.parent)
            assert res[i]["tuples"] == s.scans[0]["tuples"]
            assert np.array_equal(res[i]["W"], s.W), i
            assert np.array_equal(ctx.batch_parent(i).astype(np.uint64), s.parent), i
            assert res[i]["components"] == s.finish["components"]
            assert res[i]["largest"] == s.finish["largest"]
            got = _hash_dir(str(tmp_path / "batch" / ("sample%d" % i)))
            want = _hash_dir(str(tmp_path / ("solo%d" % i)))
            assert got == want, i
        assert nb == sum(s.out_bytes for s in solo)
        # the same batch again on the same context: identical results
        res2, _, _ = ctx.batch_partition()
        for a, b in zip(res, res2):
            assert np.array_equal(a["W"], b["W"]) and a["components"] == b["components"]


@pytest.mark.gpu
def test_batch_of_many_small_samples_equals_solo_runs():
    """150 samples of 40-400 reads of 60-150 bases: every sample is a segment
    of the batch's filter and union, so the sample boundaries put segment
    starts at nearly every bit position of the filter's 64-tuple words (a
    solo run has none inside); labels and plans equal the solo runs'."""
    print("gpu batch: 150 small samples vs solo runs", flush=True)
    rng = np.random.default_rng(9)
    bufs = [mp.synth(0, seed=900 + i, n_reads=2 * int(rng.integers(20, 201)), n_genomes=int(rng.integers(1, 4)),
                     genome_len_min=3000, genome_len_max=6000, read_len_min=60, read_len_max=150, err_rate=0.01)
            for i in range(150)]
    solo_w, solo_par, ntup = [], [], []
    with mp.Context(k=K) as ctx:
        for b in bufs:
            ctx.load(b, [0, b.nbytes])
            r = mp.partition(ctx, S=1, T=1, out_prefix=None, want_parent=True)
            solo_w.append(r.W.copy())
            solo_par.append(r.parent.astype(np.uint32))
            ntup.append(int(r.scans[0]["tuples"]))
    # the batch's segment starts: the running tuple count (the samples in order)
    starts = np.cumsum([0] + ntup[:-1]) % 64
    print("  bit positions of the segment starts: %d of 64" % len(set(starts.tolist())), flush=True)
    assert len(set(starts.tolist())) >= 48
    with mp.Context(k=K) as ctx:
        ctx.batch_load(bufs)
        res, st, _ = ctx.batch_partition()
        assert st.tuples == sum(ntup)
        for i in range(len(bufs)):
            assert np.array_equal(res[i]["W"], solo_w[i]), i
            assert np.array_equal(ctx.batch_parent(i), solo_par[i]), i
    for b in bufs:
        b.free()


@pytest.mark.gpu
def test_batch_from_device_memory_and_single_sample(tmp_path):
    print("gpu batch: device-resident input", flush=True)
    import torch
    bufs = _samples()[:3]
    host = np.concatenate([b.view() for b in bufs])
    ofs = np.concatenate([[0], np.cumsum([b.nbytes for b in bufs])]).astype(np.uint64)
    d = torch.empty(host.nbytes + 64, dtype=torch.uint8, device="cuda")
    d[: host.nbytes].copy_(torch.from_numpy(host))
    torch.cuda.synchronize()
    with mp.Context(k=K) as ctx:
        ctx.batch_load_device(d.data_ptr(), host.nbytes, ofs)
        res, _, _ = ctx.batch_partition()
        parents = [ctx.batch_parent(i) for i in range(3)]
    with mp.Context(k=K) as ctx:
        ctx.batch_load([bufs[1]])
        one, _, _ = ctx.batch_partition()
        assert np.array_equal(one[0]["W"], res[1]["W"])
        assert np.array_equal(ctx.batch_parent(0), parents[1])


@pytest.mark.gpu
def test_batch_rejects_bad_samples():
    print("gpu batch: error paths", flush=True)
    good = mp.synth(0, n_reads=2000).tobytes()
    rec = good.split(b"\n@")
    odd = b"\n@".join(rec[:-1]) + b"\n"
    with mp.Context(k=K) as ctx:
        with pytest.raises(mp.MPError) as e:
            ctx.batch_load([good, odd])
        assert e.value.code == -1
        with pytest.raises(mp.MPError) as e:
            ctx.batch_load([good] * (mp.MAX_SAMPLES + 1))
        assert e.value.code == -1
        with pytest.raises(mp.MPError) as e:
            ctx.batch_write(None)
        assert e.value.code == -7
        ctx.batch_load([good, good])
        res, _, _ = ctx.batch_partition()
        assert np.array_equal(res[0]["W"], res[1]["W"])
        assert np.array_equal(ctx.batch_parent(0), ctx.batch_parent(1))


@pytest.mark.gpu
def test_batch_with_captured_pass_chain_equals_launched(tmp_path):
    """mp_cfg.sort_graph: the batch's radix pass chains (its tuples', and the
    writer's items') are captured in HIP graphs on the first run and replayed
    on the next ones (same shape and buffers); every replay, with other work
    on the device in between, gives the labels, plans and output files of the
    pass-by-pass launch."""
    print("gpu batch: captured pass chains, 3 replays", flush=True)
    bufs = _samples()
    with mp.Context(k=K) as ctx:
        ctx.batch_load(bufs)
        want, st0, _ = ctx.batch_partition()
        assert st0.sort_graph == 0
        want_par = [ctx.batch_parent(i) for i in range(len(bufs))]
        ctx.batch_write(str(tmp_path / "eager"))
    other = mp.synth(0, n_reads=20000, seed=3)
    with mp.Context(k=K, sort_graph=True) as ctx, mp.Context(k=K) as side:
        ctx.batch_load(bufs)
        modes = []
        for rep in range(4):
            res, st, _ = ctx.batch_partition()
            modes.append(st.sort_graph)
            for a, b in zip(res, want):
                assert np.array_equal(a["W"], b["W"]) and a["components"] == b["components"], rep
            for i in range(len(bufs)):
                assert np.array_equal(ctx.batch_parent(i), want_par[i]), (rep, i)
            ctx.batch_write(str(tmp_path / ("graph%d" % rep)))
            for i in range(len(bufs)):
                assert _hash_dir(str(tmp_path / ("graph%d" % rep) / ("sample%d" % i))) == \
                    _hash_dir(str(tmp_path / "eager" / ("sample%d" % i))), (rep, i)
            side.load(other, [0, other.nbytes])  # eager work between the replays
            mp.partition(side, S=1, T=1, build_output=False)
        assert modes == [1, 2, 2, 2], modes
