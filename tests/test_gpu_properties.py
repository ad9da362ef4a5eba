"""Properties of the HIP path at sizes the oracle would take minutes for
(2 M reads of the 150 bp, k = 31 workload), determinism, and the error paths
of the C ABI (every one must return an error code, never fault)."""
import hashlib
import os

import numpy as np
import pytest

import metapart_amd as mp


def _run(buf, C=16, K=31, S=1, T=1, min3=False):
    with mp.Context(k=K, min3=min3) as ctx:
        cofs = mp.index_buffer(buf, C)
        ctx.load(buf, cofs)
        r = mp.partition(ctx, S=S, T=T, want_parent=True)
        return ctx.N, r


@pytest.mark.gpu
def test_large_run_properties_and_determinism():
    print("gpu property run: 2M reads", flush=True)
    buf = mp.synth(2, n_reads=2_000_000)
    N, r1 = _run(buf)
    assert N == 2_000_000
    sc = r1.scans[0]
    assert sc["tuples"] == int(r1.hist[-1]) == 2_000_000 * 120  # 150 bp, k = 31, no 'N'
    par = r1.parent.astype(np.int64)
    assert np.all(par >= np.arange(N))  # roots are component maxima
    assert np.all(par[par] == par)  # compressed: every parent is a root
    sizes = np.bincount(par, minlength=N)
    assert sizes.sum() == N
    W = r1.W.astype(np.int64)
    assert W[0] == 0 and W[16] == N and np.all(np.diff(W) >= 0)
    # single-end input: every record is written exactly once
    assert r1.out_bytes == buf.nbytes
    # the same input twice gives the same partition
    _, r2 = _run(buf)
    assert hashlib.sha256(r1.parent.tobytes()).hexdigest() == hashlib.sha256(r2.parent.tobytes()).hexdigest()
    assert np.array_equal(r1.W, r2.W)


@pytest.mark.gpu
def test_multiscan_equals_scan_count_invariants():
    print("gpu property run: 1M reads, S=4 T=2", flush=True)
    buf = mp.synth(2, n_reads=1_000_000, seed=9)
    N, r = _run(buf, S=4, T=2)
    assert sum(s["tuples"] for s in r.scans) == int(r.hist[-1])
    par = r.parent.astype(np.int64)
    assert np.all(par >= np.arange(N)) and np.all(par[par] == par)


@pytest.mark.gpu
def test_error_paths_return_codes():
    print("gpu error paths", flush=True)
    good = mp.synth(0, n_reads=2000).tobytes()
    with mp.Context(k=27) as ctx:
        # odd record count: loads and counts (as the reference's genHistograms
        # does), refused by the scan (the reference's pair step crashes on it)
        rec = good.split(b"\n@")
        odd = b"\n@".join(rec[:-1]) + b"\n"
        ctx.load(odd, [0, len(odd)])
        hist, _, _, _ = ctx.count()
        with pytest.raises(mp.MPError) as e:
            ctx.scan(0, 0, mp.NBUCKETS, [0, mp.NBUCKETS], 0)
        assert e.value.code == -1
        # chunk boundary inside a record
        with pytest.raises(mp.MPError) as e:
            ctx.load(good, [0, 17, len(good)])
        assert e.value.code == -1
        # quality length differs from sequence length
        bad = good.replace(b"\n+\nIIII", b"\n+\nIII", 1)
        with pytest.raises(mp.MPError) as e:
            ctx.load(bad, [0, len(bad)])
        assert e.value.code == -1
        # scan before load, finish before scan
        with mp.Context(k=27) as c2:
            with pytest.raises(mp.MPError):
                c2.scan(0, 0, 1 << 20, [0, 1 << 20])
        # a valid run still works on the same context afterwards
        ctx.load(good, [0, len(good)])
        r = mp.partition(ctx)
        assert r.W[16] == 2000
    # device budget too small: allocation is refused, not attempted
    with mp.Context(k=27, hbm_budget=1 << 16) as small:
        with pytest.raises(mp.MPError) as e:
            small.load(good, [0, len(good)])
        assert e.value.code == -2


@pytest.mark.gpu
def test_scan_labels_are_deterministic_and_compressed():
    """Lock-free union and compress: repeated runs give the same labels, and
    after a scan every non-BAD read's label is a root (compression is complete)."""
    print("gpu determinism: 3-minimizer scan, 250k reads, 4 repeats", flush=True)
    buf = mp.synth(1, n_reads=250_000, seed=3)
    seen = set()
    for rep in range(4):
        with mp.Context(k=27, min3=True) as ctx:
            ctx.load(buf, mp.index_buffer(buf, 8))
            hist, _, _, _ = ctx.count()
            scan, thread = mp.plan_ranges(hist, 1, 1)
            ctx.scan(0, int(scan[0]), int(scan[1]), thread, int(hist[int(scan[1])]))
            par = ctx.parent().astype(np.int64)
            bad = ctx.bad().astype(bool)
        ok = ~bad
        assert np.all(par[par[ok]] == par[ok])
        seen.add(hashlib.sha256(par.tobytes()).hexdigest())
    assert len(seen) == 1


@pytest.mark.gpu
def test_packing_and_writer_cover_every_read(tmp_path):
    # 4.6 M reads (more than 2^22, where an earlier wave-per-record grid had
    # to loop): the reads of the last chunk must pack like the same reads
    # loaded alone, and the writer must copy every record byte once
    print("gpu packing/writer run: 4.6M reads", flush=True)
    n, tail = 4_600_000, 50_000
    buf = mp.synth(0, n_reads=n, read_len=40, n_genomes=50, genome_len_min=20000, genome_len_max=40000, seed=31)
    v = np.frombuffer(buf.tobytes(), dtype=np.uint8)
    nl = np.flatnonzero(v == 10)
    cut = int(nl[4 * (n - tail) - 1]) + 1  # first byte of record n - tail
    with mp.Context(k=21) as ctx:
        ctx.load(buf, [0, cut, v.size])
        _, _, _, full = ctx.count(want_counts=True)
    with mp.Context(k=21) as ctx:
        ctx.load(v[cut:].copy(), [0, v.size - cut])
        _, _, _, part = ctx.count(want_counts=True)
    assert np.array_equal(full[1], part[0])
    (tmp_path / "Rank0").mkdir()
    with mp.Context(k=21) as ctx:
        ctx.load(buf, mp.index_buffer(buf, 8))
        r = mp.partition(ctx, S=1, T=1, out_prefix=str(tmp_path))
    files = sorted((tmp_path / "Rank0").iterdir())
    out = np.concatenate([np.fromfile(str(f), dtype=np.uint8) for f in files])
    assert r.out_bytes == out.size == v.size
    assert int(out.astype(np.uint64).sum()) == int(v.astype(np.uint64).sum())
    assert int(np.count_nonzero(out == ord("@"))) == int(np.count_nonzero(v == ord("@")))
