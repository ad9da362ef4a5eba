"""The CPU oracle against the real reference: every output file of the three
tools (index, KmerIdx, kmerHist, 16*T partitioned FASTQ files) must hash to
what the reference build wrote for the same inputs (tests/golden/golden.json,
made by tests/golden/make_golden.py from oracle/_ref)."""
import pytest

from _helpers import diff_hashes, golden, run_chain

CASES = [c for c in golden()["cases"]]


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_oracle_matches_reference(case, tmp_path):
    got = run_chain("oracle", case, str(tmp_path))
    assert diff_hashes(got, case["sha256"]) == []
