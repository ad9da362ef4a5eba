"""Autotuning on the device: both digit layouts of the scans' sort give the
same partition (labels and plan), and the tuners return one of the shapes or
layouts they timed."""
import numpy as np
import pytest

import metapart_amd as mp
from metapart_amd import tune


@pytest.mark.gpu
def test_digit_layouts_give_the_same_partition():
    print("gpu tune: digit layouts, 300k reads, k=31 and k=29", flush=True)
    buf = mp.synth(2, n_reads=300_000, seed=11)
    try:
        for k in (31, 29):
            out = {}
            for bits in (8, 7):
                with mp.Context(k=k, sort_digit_bits=bits) as ctx:
                    ctx.load(buf, [0, buf.nbytes])
                    r = mp.partition(ctx, S=2, T=2, want_parent=True, build_output=False)
                    out[bits] = (r.parent.copy(), r.W.copy(), [s["tuples"] for s in r.scans])
            assert np.array_equal(out[7][0], out[8][0])
            assert np.array_equal(out[7][1], out[8][1])
            assert out[7][2] == out[8][2]
    finally:
        buf.free()


@pytest.mark.gpu
def test_tuners_return_a_timed_choice():
    print("gpu tune: tuners, 200k reads", flush=True)
    d = tune.tune_sort_digits(k=31, reads=200_000)
    assert d["best"] in (7, 8) and set(d["ms_sort"]) == {7, 8}
    t = tune.tune_sort_tile(k=31, reads=200_000)
    assert t["best"] in tune.TILE_SHAPES
