"""The drop-in GPU tools (genFastqIdx, genHistograms_metapart[min],
metapart[min] over libmetapart_amd) against the real reference's output hashes,
file by file, on every golden case: plain and 3-minimizer, k = 13..32, one to
four scans, one to three segments, single-end and interleaved input, reads
with 'N', variable read lengths, reads at or below k, lower-case and other
IUPAC letters, the reference's own data/reads.fastq."""
import pytest

import os

from _helpers import diff_hashes, golden, run_chain, run_degenerate

CASES = [c for c in golden()["cases"]]
DEGENERATE = golden()["degenerate"]


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_gpu_tools_match_reference(case, tmp_path):
    print("gpu golden case", case["name"], flush=True)
    got = run_chain("gpu", case, str(tmp_path), timeout=300)
    assert diff_hashes(got, case["sha256"]) == []


STATUS64 = ["c1_T2", "gcrep_k31_S4T2", "mixed_k27_S2T2", "var_k13_S2T2"]


@pytest.mark.gpu
@pytest.mark.parametrize("name", STATUS64)
def test_gpu_tools_match_reference_with_64bit_status(name, tmp_path, monkeypatch):
    """The same with the radix passes' 64-bit look-back status words forced."""
    print("gpu golden case (64-bit status)", name, flush=True)
    monkeypatch.setenv("MP_SORT_STATUS64", "1")
    monkeypatch.setenv("MP_SORT_LOOKBACK", "1")  # every pass by look-back, none from digit counts
    case = next(c for c in CASES if c["name"] == name)
    got = run_chain("gpu", case, str(tmp_path), timeout=300)
    assert diff_hashes(got, case["sha256"]) == []


BYTE_DIGITS = ["c1_S3T2_k31", "small_k31_S2T2", "gcrep_k31_S4T2", "il_k31_T2", "mixed_k31", "small_k30_S2T2"]


@pytest.mark.gpu
@pytest.mark.parametrize("name", BYTE_DIGITS)
def test_gpu_tools_match_reference_with_byte_digits(name, tmp_path, monkeypatch):
    """The same with every radix pass on a whole byte of the k-mer (at k = 31
    the default splits the top 14 bits over the last two passes, 7 + 7)."""
    print("gpu golden case (byte digits)", name, flush=True)
    monkeypatch.setenv("MP_SORT_DIGITS", "8")
    case = next(c for c in CASES if c["name"] == name)
    got = run_chain("gpu", case, str(tmp_path), timeout=300)
    assert diff_hashes(got, case["sha256"]) == []


# Sort shapes a caller or the autotuner can choose (mp_cfg.sort_tile: 512 x 6,
# 512 x 8, 512 x 12 and 1024 x 8 tuples a tile), tiles taken in launch order
# instead of one contiguous range per XCD, and every pass global (no in-place
# local pass), each on golden cases it does not otherwise meet.
SORT_VARIANTS = [("MP_SORT_TILE", "3072", "rnd6_k27_S3T4"), ("MP_SORT_TILE", "4096", "rndL3_k21_S1T1"),
                 ("MP_SORT_TILE", "6144", "c1_S3T2_k31"), ("MP_SORT_TILE", "8192", "gcrep_k31_S4T2"),
                 ("MP_SORT_TILE", "4096", "rndL5_k24_S1T2_min3"), ("MP_SORT_XCD", "0", "small_k31_S2T2"),
                 ("MP_SORT_GLOBAL", "1", "mixed_k31"), ("MP_SORT_GLOBAL", "1", "rnd7_k29_S3T2")]


@pytest.mark.gpu
@pytest.mark.parametrize("var", SORT_VARIANTS, ids=["%s=%s-%s" % v for v in SORT_VARIANTS])
def test_gpu_tools_match_reference_with_sort_variants(var, tmp_path, monkeypatch):
    env, val, name = var
    print("gpu golden case (%s=%s)" % (env, val), name, flush=True)
    monkeypatch.setenv(env, val)
    case = next(c for c in CASES if c["name"] == name)
    got = run_chain("gpu", case, str(tmp_path), timeout=300)
    assert diff_hashes(got, case["sha256"]) == []


@pytest.mark.gpu
@pytest.mark.parametrize("case", DEGENERATE, ids=[c["name"] for c in DEGENERATE])
def test_gpu_tools_on_input_the_reference_cannot_partition(case, tmp_path):
    """No read holds a window of k, or the read count is odd: the index and
    the histograms equal the reference's; metapart, where the reference
    crashes, exits with an error and writes no output."""
    print("gpu degenerate case", case["name"], flush=True)
    got, r = run_degenerate("gpu", case, str(tmp_path))
    assert diff_hashes(got, case["sha256"]) == []
    assert r.returncode != 0 and case["refusal"] in r.stdout
    assert os.listdir(os.path.join(str(tmp_path), "out", "Rank0")) == []
