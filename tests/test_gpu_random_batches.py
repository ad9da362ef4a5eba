"""Batches drawn from a fixed seed: k = 15..31, plain and 3-minimizer, 2-40
samples of 2-3000 reads of varied lengths, error rates and letters.  Every
sample's labels, plan and component counts in the batch equal its solo run,
and three samples of each batch equal the oracle's labels."""
import os

import numpy as np
import pytest

from _helpers import oracle_run
from test_gpu_stages import _mix_letters

import metapart_amd as mp


def _batches(n=6, seed=4242):
    rng = np.random.default_rng(seed)
    out = []
    for b in range(n):
        min3 = bool(rng.random() < 0.3)
        K = int(rng.integers(15, 32))
        ns = int(rng.integers(2, 41))
        samples = []
        for i in range(ns):
            lo = int(rng.integers(K + 5, 130))
            over = dict(seed=7000 + 100 * b + i, n_reads=2 * int(rng.integers(1, 1501)), n_genomes=int(rng.integers(1, 8)),
                        genome_len_min=2000, genome_len_max=12000, read_len_min=lo, read_len_max=int(rng.integers(lo, 200)),
                        err_rate=float(rng.choice([0.0, 0.005, 0.01, 0.02])), n_rate=float(rng.choice([0.0, 0.002])))
            if rng.random() < 0.2:
                over["mix"] = 0.05
            samples.append(over)
        out.append(("b%d_k%d%s_n%d" % (b, K, "_min3" if min3 else "", ns), K, min3, samples))
    return out


def _make(over):
    over = dict(over)
    mix = over.pop("mix", 0.0)
    buf = mp.synth(0, **over)
    return _mix_letters(buf, mix, over["seed"]) if mix else np.frombuffer(buf.tobytes(), dtype=np.uint8).copy()


CASES = _batches()


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_random_batch_equals_solo_runs_and_oracle(case, tmp_path):
    name, K, min3, samples = case
    print("gpu random batch", name, flush=True)
    bufs = [_make(o) for o in samples]
    solo = []
    with mp.Context(k=K, min3=min3) as ctx:
        for b in bufs:
            ctx.load(b, [0, b.nbytes])
            r = mp.partition(ctx, S=1, T=1, out_prefix=None, want_parent=True)
            solo.append((r.W.copy(), r.parent.astype(np.uint32), r.finish["components"]))
    with mp.Context(k=K, min3=min3) as ctx:
        ctx.batch_load(bufs)
        res, _, _ = ctx.batch_partition()
        for i in range(len(bufs)):
            assert np.array_equal(res[i]["W"], solo[i][0]), i
            assert np.array_equal(ctx.batch_parent(i), solo[i][1]), i
            assert res[i]["components"] == solo[i][2], i
    for i in sorted({0, len(bufs) // 2, len(bufs) - 1}):
        p = str(tmp_path / ("s%d.fastq" % i))
        open(p, "wb").write(bufs[i].tobytes())
        ref = oracle_run(p, 1, K, 1, 1, min3=min3, out_dir=str(tmp_path / ("o%d" % i)))
        assert np.array_equal(solo[i][1], ref["parent"].astype(np.uint32)), i
