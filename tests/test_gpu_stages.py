"""Stage-by-stage parity of the HIP path (through the C ABI) with the oracle on
the same seeded inputs: per-chunk m-mer counts, the filter's evaluated and
zeroed run counts of every scan, the sorted and filtered tuple array of a scan
(reference 14-byte layout), BAD flags and parent labels after that scan, and
the final parent labels after the pair step."""
import os

import numpy as np
import pytest

from _helpers import BAD, oracle_run, read_kmeridx

import metapart_amd as mp

CASES = [
    # name, synth overrides, C, K, S, T, min3, interleaved, dump_scan
    ("plain_k27", dict(n_reads=6000, seed=41), 4, 27, 1, 1, False, False, 0),
    ("plain_k31_T2", dict(n_reads=6000, seed=42, read_len=150), 4, 31, 1, 2, False, False, 0),
    ("plain_k31_S3_scan2", dict(n_reads=8000, seed=43, read_len=150), 6, 31, 3, 2, False, False, 2),
    ("plain_k20_S2_scan1", dict(n_reads=8000, seed=44, n_rate=0.003), 4, 20, 2, 1, False, False, 1),
    ("min3_k27", dict(n_reads=8000, seed=45), 4, 27, 1, 1, True, False, 0),
    ("il_k31", dict(n_reads=8000, seed=46, read_len=150, interleaved=1), 4, 31, 1, 1, False, True, 0),
    ("varlen_k25_T3", dict(n_reads=9000, seed=47, read_len_min=40, read_len_max=180), 6, 25, 1, 3, False, False, 0),
    # reads of 22-30 bases at K = 27: most have no k-mer or a few, so a sort
    # tile's windows span thousands of reads (more than its LDS read table)
    ("short_k27", dict(n_reads=30000, seed=48, read_len_min=22, read_len_max=30, n_rate=0.002), 4, 27, 1, 1, False,
     False, 0),
    # 40 small genomes: components of a few hundred reads that the pair step
    # chains past 1000 reads, so part of it runs in the reference's order
    ("pairs_ordered", dict(n_reads=20000, n_genomes=40, genome_len_min=4000, genome_len_max=8000, abund_sigma=1.0,
                           seed=8), 4, 31, 1, 1, False, False, 0),
    # many thread segments: segment starts inside the filter's tiles and words,
    # where a run of one tuple ends its segment and a run's end is a segment's
    ("plain_k31_T16", dict(n_reads=12000, seed=49, read_len=150), 16, 31, 1, 16, False, False, 0),
    ("plain_k25_S2T7", dict(n_reads=10000, seed=50), 14, 25, 2, 7, False, False, 1),
]
ORDERED = {"pairs_ordered"}
CASES.append(
    # one 3 kb genome at ~3000x: every k-mer has thousands of copies, so the
    # sort's local pass hands whole groups to the per-group global passes
    ("deep_repeat", dict(n_reads=60000, n_genomes=1, genome_len_min=3000, genome_len_max=3000, read_len=150,
                         err_rate=0.005, seed=12), 4, 31, 1, 1, False, False, 0))
CASES.append(
    # one 400 bp genome at ~60000x: k-mer runs of up to 72 k tuples (past the
    # filter's 16-bit tallies of a staged run, so evaluated from memory) and
    # sort groups far past the local pass's window
    ("ultra_deep", dict(n_reads=160000, n_genomes=1, genome_len_min=400, genome_len_max=400, read_len=150,
                        err_rate=0.002, seed=13), 4, 31, 1, 1, False, False, 0))
LONG_GROUPS = {"deep_repeat", "ultra_deep"}
LONG_RUNS = {"ultra_deep"}


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_stages_match_oracle(case, tmp_path):
    print("gpu stage case", case[0], flush=True)
    _check_stages(case, tmp_path)


# The radix passes that find their offsets by look-back (the generating pass;
# every pass with MP_SORT_LOOKBACK) switch to 64-bit status words when a digit
# of a pass holds 2^26 tuples or more; MP_SORT_STATUS64 forces that format, so
# it meets the oracle on small inputs.
STATUS64 = ["plain_k31_T2", "plain_k31_S3_scan2", "varlen_k25_T3", "pairs_ordered", "deep_repeat"]


@pytest.mark.gpu
@pytest.mark.parametrize("name", STATUS64)
def test_stages_match_oracle_with_64bit_status(name, tmp_path, monkeypatch):
    print("gpu stage case (64-bit status)", name, flush=True)
    monkeypatch.setenv("MP_SORT_STATUS64", "1")
    monkeypatch.setenv("MP_SORT_LOOKBACK", "1")  # every pass by look-back, none from digit counts
    _check_stages(next(c for c in CASES if c[0] == name), tmp_path)


def _mix_letters(buf, rate, seed):
    """The FASTQ bytes with a share `rate` of the sequence bases changed: nine
    in ten to lower case, the rest to another IUPAC letter or 'n' (bytes whose
    2-bit code, adj byte and 'N' test the reference takes from the raw
    character)."""
    a = np.frombuffer(buf.tobytes(), dtype=np.uint8).copy()
    nl = np.flatnonzero(a == ord("\n"))
    starts, ends = nl[0::4] + 1, nl[1::4]  # sequence lines: after the 1st newline of a record, to the 2nd
    inseq = np.zeros(a.size + 1, dtype=np.int32)
    np.add.at(inseq, starts, 1)
    np.add.at(inseq, ends, -1)
    inseq = np.cumsum(inseq[:-1]) > 0
    rng = np.random.default_rng(seed)
    pick = inseq & (rng.random(a.size) < rate)
    low = pick & (rng.random(a.size) < 0.9)
    a[low] |= 0x20
    other = np.frombuffer(b"RYKMSWBDHVn", dtype=np.uint8)
    odd = pick & ~low
    a[odd] = other[rng.integers(0, other.size, int(odd.sum()))]
    return a


def _check_stages(case, tmp_path):
    name, over, C, K, S, T, min3, il, dump = case
    over = dict(over)
    mix = over.pop("mix", 0.0)
    buf = mp.synth(0, **over)
    if mix:
        buf = _mix_letters(buf, mix, over.get("seed", 0))
    p = os.path.join(str(tmp_path), "in.fastq")
    open(p, "wb").write(buf.tobytes())
    ref = oracle_run(p, C, K, S, T, interleaved=il, min3=min3, dump_scan=dump)
    _, ref_counts = read_kmeridx(p + (".KmerIdx_3min%d" % K if min3 else ".KmerIdx%d" % K))

    with mp.Context(k=K, min3=min3) as ctx:
        cofs = mp.index_buffer(buf, C, interleaved=il)
        ctx.load(buf, cofs)
        hist, reads, tot, counts = ctx.count(want_counts=True)
        assert np.array_equal(counts, ref_counts[:-1])
        assert int(hist[-1]) == tot == int(ref_counts.sum())
        scan, thread = mp.plan_ranges(hist, S, T)
        for s in range(S):
            lo, hi = int(scan[s]), int(scan[s + 1])
            st = ctx.scan(s, lo, hi, thread[s * T:(s + 1) * T + 1], int(hist[min(hi, mp.NBUCKETS)] - hist[lo]))
            assert (st.evaluated_runs, st.zeroed_runs) == ref["filter_counts"][s]
            if name in LONG_GROUPS:
                assert st.sort_long_groups > 0
            if s == dump:
                if name in LONG_RUNS:  # the case holds the runs it is meant to
                    keys = ref["tuples"][:, :8].copy().view(np.uint64).ravel()
                    starts = np.concatenate([[0], np.flatnonzero(np.diff(keys)) + 1, [keys.size]])
                    assert np.diff(starts).max() > 65535
                tup = ctx.dump_tuples()
                assert tup.shape == ref["tuples"].shape
                assert np.array_equal(tup, ref["tuples"])
                par = ctx.parent().astype(np.uint64)
                bad = ctx.bad().astype(bool)
                ref_par = ref["parent_scan"]
                ref_bad = ref_par == BAD
                assert np.array_equal(bad, ref_bad)
                assert np.array_equal(par[~bad], ref_par[~ref_bad])
                assert np.array_equal(par[bad], np.nonzero(bad)[0].astype(np.uint64))
        W, fs = ctx.finish()
        assert np.array_equal(ctx.parent(), ref["parent"])
        if name in ORDERED:
            assert fs.pair_candidates > 0
        assert W[0] == 0 and W[16] == ctx.N and np.all(np.diff(W.astype(np.int64)) >= 0)
