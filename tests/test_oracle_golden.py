"""The CPU oracle against the real reference: every output file of the three
tools (index, KmerIdx, kmerHist, 16*T partitioned FASTQ files) must hash to
what the reference build wrote for the same inputs (tests/golden/golden.json,
made by tests/golden/make_golden.py from oracle/_ref)."""
import os

import pytest

from _helpers import diff_hashes, golden, golden_large, make_large_input, run_chain, run_degenerate

CASES = [c for c in golden()["cases"]]


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_oracle_matches_reference(case, tmp_path):
    got = run_chain("oracle", case, str(tmp_path))
    assert diff_hashes(got, case["sha256"]) == []


DEGENERATE = golden()["degenerate"]


@pytest.mark.parametrize("case", DEGENERATE, ids=[c["name"] for c in DEGENERATE])
def test_oracle_on_input_the_reference_cannot_partition(case, tmp_path):
    """No read holds a window of k, or the read count is odd: the index and
    the histograms equal the reference's; metapart, where the reference
    crashes (it walks past its range arrays on an empty histogram; its pair
    step reads id N of an odd count), fails with an error and writes no
    output."""
    assert case["reference_metapart_status"] != 0
    got, r = run_degenerate("oracle", case, str(tmp_path))
    assert diff_hashes(got, case["sha256"]) == []
    assert r.returncode != 0 and case["refusal"] in r.stdout
    assert os.listdir(os.path.join(str(tmp_path), "out", "Rank0")) == []


# Workload-size cases (golden_large.json): minutes of CPU time and up to ~60 GB
# of host memory each, so they run only when MP_SLOW_ORACLE=1.  Checked this
# way: c3slice_k31_S4 (2 M reads, 64 chunks, four scans) and c2_min3_k31
# (20 M reads, 3-minimizer); the two-scan 20 M-read plain cases need three
# 16.8 GB tuple buffers at once in this restatement.
LARGE = golden_large()["cases"]


@pytest.mark.skipif(os.environ.get("MP_SLOW_ORACLE") != "1", reason="minutes of CPU per case; MP_SLOW_ORACLE=1 runs it")
@pytest.mark.parametrize("case", LARGE, ids=[c["name"] for c in LARGE])
def test_oracle_matches_reference_at_workload_size(case, tmp_path):
    src = str(tmp_path / "src.fastq")
    make_large_input(case["input"], src)
    got = run_chain("oracle", case, str(tmp_path / "run"), timeout=7200, input_path=src)
    assert diff_hashes(got, case["sha256"]) == []
