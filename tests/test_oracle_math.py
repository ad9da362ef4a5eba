"""Pins of the CPU oracle by what the mathematics fixes, independent of the
reference run: canonical k-mers by string reverse-complement, 'N' windows,
m-mer buckets, tuple strand and neighbour bytes, sort order, and the union as
connected components whose root is the largest read id.  (The frequency
filter, the 3-minimizer windows and the pair step have no closed form: their
pin is the reference run in test_oracle_golden.py.)"""
import ctypes as ct
import os
import random

import numpy as np
import pytest

from _helpers import BAD, ROOT, oracle, oracle_run

CODE = {"A": 0, "C": 1, "G": 2, "T": 3}
COMP = {"A": "T", "C": "G", "G": "C", "T": "A", "N": "N"}
# bits of the ASCII codes the neighbour byte is built from: low field (c>>1)&7
# and high field ((c<<2)&0xF8)>>3 for the five letters the data uses
LOW = {"A": 0, "C": 1, "G": 3, "T": 2, "N": 7}
HIGH = {"A": 0, "C": 1, "G": 3, "T": 10, "N": 7}


def num(s):
    v = 0
    for ch in s:
        v = v * 4 + CODE[ch]
    return v


def revcomp(s):
    return "".join(COMP[c] for c in reversed(s))


def canonical_windows(seq, K):
    out = []
    for p in range(len(seq) - K + 1):
        w = seq[p:p + K]
        if "N" in w:
            out.append(None)
            continue
        f, r = num(w), num(revcomp(w))
        out.append((min(f, r), f < r, p))
    return out


def rand_read(rng, L, nrate=0.0):
    return "".join("N" if rng.random() < nrate else rng.choice("ACGT") for _ in range(L))


@pytest.mark.parametrize("K", [11, 20, 27, 31])
def test_count_matches_canonical_bruteforce(K):
    rng = random.Random(K)
    L = oracle()
    for trial in range(40):
        Lr = rng.randint(K - 2, K + 90)
        seq = rand_read(rng, Lr, 0.02 if trial % 3 == 0 else 0.0)
        counts = np.zeros(2 ** 20 + 2, dtype=np.uint32)
        n = L.orc_count_read(seq.encode(), len(seq), K, 0, counts.ctypes.data)
        expect = np.zeros_like(counts)
        m = 0
        if len(seq) > K:  # a read of exactly K bases yields nothing (countMmers.c:87)
            for w in canonical_windows(seq, K):
                if w is not None:
                    expect[w[0] >> (2 * (K - 10))] += 1
                    m += 1
        assert n == m
        assert np.array_equal(counts, expect)


@pytest.mark.parametrize("K", [21, 31])
def test_enum_tuples_bruteforce(K):
    """Tuple order = position order; k-mer = canonical; adj = the neighbours on
    the chosen strand ('N' outside the read)."""
    rng = random.Random(100 + K)
    L = oracle()
    for trial in range(30):
        seq = rand_read(rng, rng.randint(K + 1, K + 80), 0.01)
        out = np.zeros((300, 14), dtype=np.uint8)
        label = rng.randint(0, 10 ** 6)
        n = L.orc_enum_read(seq.encode(), len(seq), K, 0, 0, 2 ** 20, label, out.ctypes.data, 300)
        exp = []
        for w in canonical_windows(seq, K):
            if w is None:
                continue
            c, fwd, p = w
            prev = seq[p - 1] if p > 0 else "N"
            nxt = seq[p + K] if p + K < len(seq) else "N"
            if fwd:
                before, after = prev, nxt
            else:  # on the reverse strand the successor is comp(prev), the predecessor comp(next)
                before, after = COMP[nxt], COMP[prev]
            exp.append((c, (HIGH[before] << 3) | LOW[after]))
        assert n == len(exp)
        for i, (c, a) in enumerate(exp):
            assert int(out[i, :8].view(np.uint64)[0]) == c
            assert out[i, 8] == (a & 0xFF)
            assert int(out[i, 10:14].view(np.uint32)[0]) == label


def test_enum_range_filter():
    rng = random.Random(7)
    L = oracle()
    K = 27
    seq = rand_read(rng, 150)
    full = np.zeros((300, 14), dtype=np.uint8)
    n = L.orc_enum_read(seq.encode(), len(seq), K, 0, 0, 2 ** 20, 5, full.ctypes.data, 300)
    buckets = [int(full[i, :8].view(np.uint64)[0]) >> (2 * (K - 10)) for i in range(n)]
    lo, hi = sorted(buckets)[n // 3], sorted(buckets)[2 * n // 3]
    part = np.zeros((300, 14), dtype=np.uint8)
    m = L.orc_enum_read(seq.encode(), len(seq), K, 0, lo, hi, 5, part.ctypes.data, 300)
    keep = [full[i] for i in range(n) if lo <= buckets[i] < hi]
    assert m == len(keep)
    assert all(np.array_equal(part[i], keep[i]) for i in range(m))


def test_min3_emits_canonical_kmers_of_the_read():
    rng = random.Random(9)
    L = oracle()
    K = 27
    for _ in range(50):
        seq = rand_read(rng, rng.randint(K + 1, 160), 0.01)
        cset = {w[0] for w in canonical_windows(seq, K) if w is not None}
        out = np.zeros((300, 14), dtype=np.uint8)
        n = L.orc_enum_read(seq.encode(), len(seq), K, 1, 0, 2 ** 20, 1, out.ctypes.data, 300)
        assert n <= max(1, (len(seq) - K)) + 1
        for i in range(n):
            k = int(out[i, :8].view(np.uint64)[0])
            assert k in cset and k != 0


def test_plan_ranges_properties():
    rng = np.random.default_rng(3)
    L = oracle()
    counts = rng.integers(0, 50, size=2 ** 20).astype(np.uint64)
    H = np.zeros(2 ** 20 + 1, dtype=np.uint64)
    H[1:] = np.cumsum(counts)
    for S, T in [(1, 1), (1, 4), (3, 2), (5, 3)]:
        scan = np.zeros(S + 1, dtype=np.uint64)
        thr = np.zeros(S * T + 1, dtype=np.uint64)
        assert L.orc_plan_ranges(H.ctypes.data, S, T, scan.ctypes.data, thr.ctypes.data) == 0
        assert np.all(np.diff(thr.astype(np.int64)) > 0)
        assert np.array_equal(thr[::T], scan)
        # every scan boundary is the first bucket whose prefix reaches r * ceil(tot / S)
        tot = int(H[-1])
        avg = -(-tot // S)
        for r in range(1, S):
            assert int(scan[r]) == int(np.argmax(H >= avg * r))
        assert int(H[int(scan[S])]) == tot or int(scan[S]) == 2 ** 20


def _components(tuples, bad, K):
    """Connected components of reads joined by runs of equal non-zero k-mers,
    BAD labels excluded; returns expected root (max id) per read."""
    keys = tuples[:, :8].copy().view(np.uint64).ravel()
    labs = tuples[:, 10:14].copy().view(np.uint32).ravel()
    N = len(bad)
    adj = [[] for _ in range(N)]
    i = 0
    while i < len(keys):
        j = i
        while j < len(keys) and keys[j] == keys[i]:
            j += 1
        if keys[i] != 0:
            alive = [int(x) for x in labs[i:j] if not bad[x]]
            for x in alive[1:]:
                adj[alive[0]].append(x)
                adj[x].append(alive[0])
        i = j
    root = np.arange(N, dtype=np.int64)
    seen = np.zeros(N, dtype=bool)
    for s in range(N):
        if seen[s] or bad[s]:
            continue
        comp, stack = [], [s]
        seen[s] = True
        while stack:
            x = stack.pop()
            comp.append(x)
            for y in adj[x]:
                if not seen[y]:
                    seen[y] = True
                    stack.append(y)
        m = max(comp)
        for x in comp:
            root[x] = m
    return root


@pytest.mark.parametrize("K,T", [(27, 1), (31, 1)])
def test_union_is_connected_components_with_max_root(tmp_path, K, T):
    import metapart_amd as mp
    buf = mp.synth(0, n_reads=3000, seed=31 + K)
    p = os.path.join(str(tmp_path), "in.fastq")
    open(p, "wb").write(buf.tobytes())
    res = oracle_run(p, 4, K, 1, T, dump_scan=0)
    tup, ps = res["tuples"], res["parent_scan"]
    keys = tup[:, :8].copy().view(np.uint64).ravel()
    nz = keys[keys != 0]
    assert np.all(nz[1:] >= nz[:-1])  # sorted by k-mer (zeroed runs keep their place)
    bad = ps == BAD
    exp = _components(tup, bad, K)
    assert np.array_equal(ps[~bad].astype(np.int64), exp[~bad])


def test_sort_is_stable_by_kmer_then_adj(tmp_path):
    import metapart_amd as mp
    buf = mp.synth(0, n_reads=2000, seed=77)
    p = os.path.join(str(tmp_path), "in.fastq")
    open(p, "wb").write(buf.tobytes())
    res = oracle_run(p, 4, 27, 1, 1, dump_scan=0)
    tup = res["tuples"]
    keys = tup[:, :8].copy().view(np.uint64).ravel()
    adj = tup[:, 8].astype(np.int64)
    labs = tup[:, 10:14].copy().view(np.uint32).ravel().astype(np.int64)
    nz = np.nonzero(keys != 0)[0]
    k2, a2, l2 = keys[nz], adj[nz], labs[nz]
    same_k = k2[1:] == k2[:-1]
    assert np.all(a2[1:][same_k] >= a2[:-1][same_k])
    same = same_k & (a2[1:] == a2[:-1])
    assert np.all(l2[1:][same] >= l2[:-1][same])  # scan 0: generation order is read-id order
