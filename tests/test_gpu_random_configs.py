"""Stage-by-stage parity with the oracle over 24 configurations drawn from a
fixed seed: k = 13..32, one to three scans, one to eight thread segments
(dividing the chunk count), plain and 3-minimizer, single-end and
interleaved, read lengths from 15 to 200 bases, 'N' rates up to 0.5 %, and in a
quarter of them 5 % of the bases lower-case or other IUPAC letters.  The
fixed cases elsewhere pin chosen corners; these put segment starts, run
boundaries and tile ends at positions nobody chose.  The 3-minimizer variant
and k = 32 take one scan (the reference's plans for more are degenerate)."""
import numpy as np
import pytest

from test_gpu_stages import _check_stages


def _configs(n=24, seed=2027):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        min3 = bool(rng.random() < 0.3)
        K = int(rng.integers(13, 32 if min3 else 33))
        C = int(rng.choice([4, 6, 8, 12, 16]))
        T = int(rng.choice([t for t in (1, 2, 3, 4, 6, 8) if C % t == 0]))
        S = int(rng.integers(1, 4))
        if min3 or K == 32:
            S = 1
        il = bool(rng.random() < 0.2)
        lo = int(rng.integers(max(K - 5, 15), 120))
        hi = int(rng.integers(lo, 200))
        over = dict(n_reads=2 * int(rng.integers(1500, 4001)), seed=3000 + i, read_len_min=lo, read_len_max=hi,
                    n_rate=float(rng.choice([0.0, 0.001, 0.005])), err_rate=float(rng.choice([0.005, 0.01, 0.02])),
                    n_genomes=int(rng.integers(2, 20)), genome_len_min=3000, genome_len_max=20000)
        if il:
            over["interleaved"] = 1
        if rng.random() < 0.25:  # lower-case and other IUPAC letters in 5 % of the bases
            over["mix"] = 0.05
        dump = int(rng.integers(0, S))
        name = "r%02d_k%d_S%dT%d%s%s%s" % (i, K, S, T, "_min3" if min3 else "", "_il" if il else "",
                                          "_mix" if "mix" in over else "")
        out.append((name, over, C, K, S, T, min3, il, dump))
    return out


CASES = _configs()


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_random_config_stages_match_oracle(case, tmp_path):
    print("gpu random case", case[0], flush=True)
    _check_stages(case, tmp_path)
