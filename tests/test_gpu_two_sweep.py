"""Scans that do not fit the device budget go twice over bucket parts (a
filter sweep that completes the BAD set, then a union sweep).  Every result
must equal the one-pass scan of the same reads: labels, plan, partition files,
and the oracle's labels."""
import hashlib
import os

import numpy as np
import pytest

from _helpers import oracle_run

import metapart_amd as mp

CASES = [
    # name, synth overrides, C, K, S, T, min3, interleaved
    ("plain_k31_T2", dict(n_reads=20000, seed=61, read_len=150), 4, 31, 1, 2, False, False),
    ("plain_k25_S2_T3", dict(n_reads=24000, seed=62, read_len_min=50, read_len_max=160, n_rate=0.002), 6, 25, 2, 3,
     False, False),
    ("il_k31_T2", dict(n_reads=20000, seed=64, read_len=150, interleaved=1), 4, 31, 1, 2, False, True),
    # eight thread segments: the parts begin and end inside segments
    ("plain_k27_T8", dict(n_reads=24000, seed=65, read_len_min=60, read_len_max=150), 8, 27, 1, 8, False, False),
]


def _files(d):
    out = {}
    for f in sorted(os.listdir(os.path.join(d, "Rank0"))):
        out[f] = hashlib.sha256(open(os.path.join(d, "Rank0", f), "rb").read()).hexdigest()
    return out


def _run(buf, C, K, S, T, min3, il, out, budget=0):
    os.makedirs(os.path.join(out, "Rank0"), exist_ok=True)
    with mp.Context(k=K, min3=min3, hbm_budget=budget) as ctx:
        cofs = mp.index_buffer(buf, C, interleaved=il)
        ctx.load(buf, cofs)
        il_start = None
        r = mp.partition(ctx, S=S, T=T, out_prefix=out, want_parent=True, il_start=il_start)
        return r, ctx.device_bytes


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_two_sweeps_equal_one_pass(case, tmp_path):
    name, over, C, K, S, T, min3, il = case
    print("gpu two-sweep case", name, flush=True)
    buf = mp.synth(0, **over)
    one, _ = _run(buf, C, K, S, T, min3, il, str(tmp_path / "one"))
    big = max(int(sc["tuples"]) for sc in one.scans)
    os.environ["MP_SCAN_PART_TUPLES"] = str(big // 3 + 1)
    try:
        two, _ = _run(buf, C, K, S, T, min3, il, str(tmp_path / "two"))
    finally:
        del os.environ["MP_SCAN_PART_TUPLES"]
    assert max(int(sc["parts"]) for sc in two.scans) >= 3
    assert all(int(sc["parts"]) == 0 for sc in one.scans)
    for a, b in zip(one.scans, two.scans):
        for f in ("tuples", "bad_labels", "zeroed_runs", "evaluated_runs", "quirk_segments"):
            assert a[f] == b[f], f
    assert np.array_equal(one.parent, two.parent)
    assert np.array_equal(one.W, two.W)
    assert _files(str(tmp_path / "one")) == _files(str(tmp_path / "two"))
    p = str(tmp_path / "in.fastq")
    open(p, "wb").write(buf.tobytes())
    ref = oracle_run(p, C, K, S, T, interleaved=il, min3=min3)
    assert np.array_equal(two.parent, ref["parent"])


@pytest.mark.gpu
def test_three_minimizer_scan_that_does_not_fit_fails(tmp_path):
    # a 3-minimizer enumeration over a bucket sub-range is not the histogram's
    # share of it, so such a scan is not split: it fails with ENOMEM
    print("gpu two-sweep min3 refusal", flush=True)
    buf = mp.synth(0, n_reads=20000, seed=63)
    one, _ = _run(buf, 4, 27, 1, 1, True, False, str(tmp_path / "one"))
    os.environ["MP_SCAN_PART_TUPLES"] = str(int(one.scans[0]["tuples"]) // 2)
    try:
        with pytest.raises(mp.MPError) as e:
            _run(buf, 4, 27, 1, 1, True, False, str(tmp_path / "two"))
        assert e.value.code == -2
    finally:
        del os.environ["MP_SCAN_PART_TUPLES"]


@pytest.mark.gpu
def test_two_sweeps_from_the_budget(tmp_path):
    # no override: a budget that holds the reads but not one scan's tuples
    print("gpu two-sweep budget case", flush=True)
    buf = mp.synth(0, n_reads=40000, seed=65, read_len=150)
    one, dev = _run(buf, 4, 31, 1, 1, False, False, str(tmp_path / "one"))
    n = int(one.scans[0]["tuples"])
    two, _ = _run(buf, 4, 31, 1, 1, False, False, str(tmp_path / "two"), budget=dev - 14 * n)
    assert int(two.scans[0]["parts"]) >= 2
    assert np.array_equal(one.parent, two.parent)
    assert np.array_equal(one.W, two.W)
    assert _files(str(tmp_path / "one")) == _files(str(tmp_path / "two"))
