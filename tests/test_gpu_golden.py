"""The drop-in GPU tools (genFastqIdx, genHistograms_metapart[min],
metapart[min] over libmetapart_amd) against the real reference's output hashes,
file by file, on every golden case: plain and 3-minimizer, k = 13..32, one to
four scans, one to three segments, single-end and interleaved input, reads
with 'N', variable read lengths, reads at or below k, lower-case and other
IUPAC letters, the reference's own data/reads.fastq."""
import pytest

from _helpers import diff_hashes, golden, run_chain

CASES = [c for c in golden()["cases"]]


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_gpu_tools_match_reference(case, tmp_path):
    print("gpu golden case", case["name"], flush=True)
    got = run_chain("gpu", case, str(tmp_path), timeout=300)
    assert diff_hashes(got, case["sha256"]) == []
