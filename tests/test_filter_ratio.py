"""The filter's ratio test on a run held in its tile (at most 6400 tuples):
(double)((float)smax / (float)deg) > 0.2 (metapart.c:2356-2359, 2419) as the
integer comparison smax * 2^27 > 26843545 * deg that the kernel evaluates
(mp_graph.hip, staged_verdict).  Checked for every smax <= deg <= 6400 against
numpy's IEEE fp32 division."""
import numpy as np

MAX_RUN = 6400  # the filter tile's staged tuples


def test_integer_ratio_equals_fp32_division():
    s = np.arange(MAX_RUN + 1, dtype=np.int64)
    for deg in range(1, MAX_RUN + 1):
        ss = s[: deg + 1]
        ref = (ss.astype(np.float32) / np.float32(deg)).astype(np.float64) > 0.2
        got = (ss << 27) > 26843545 * deg
        assert np.array_equal(ref, got), deg
