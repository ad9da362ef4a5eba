"""The C ABI library without a GPU: it loads, exports every entry point the
header declares, and its host-side parts (chunk indexer, range planner) match
the reference (golden hashes) and the oracle."""
import ctypes as ct
import os
import re
import subprocess

import numpy as np
import pytest

from _helpers import BIN, ORACLE_CLI, ROOT, golden, make_input, masked_sha, oracle

import metapart_amd as mp


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "metapart_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mp_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = ct.CDLL(mp.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 18
    missing = [n for n in names if not hasattr(L, n)]
    assert missing == []
    assert sorted(mp.SYMBOLS) == names


def test_ctx_create_without_gpu_fails_cleanly():
    # on a machine without a GPU the library reports an error instead of aborting
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    with pytest.raises(mp.MPError):
        mp.Context(k=31)


def test_ctx_create_rejects_unsupported_sort_shapes():
    # checked before any device call: an unsupported digit width or tile shape is MP_EINVAL on any machine
    import ctypes as ct
    for kw in (dict(sort_digit_bits=11), dict(sort_digit_bits=4), dict(sort_tile=5000)):
        cfg = mp.Cfg(device=0, k=31, min3=0, hbm_budget=0, sort_digit_bits=0, sort_tile=0)
        for f, v in kw.items():
            setattr(cfg, f, v)
        h = ct.c_void_p()
        assert mp.lib().mp_ctx_create(ct.byref(cfg), ct.byref(h)) == -1
        assert not h.value


@pytest.mark.parametrize("case", [c for c in golden()["cases"]], ids=[c["name"] for c in golden()["cases"]])
def test_index_tool_matches_reference(case, tmp_path):
    d = str(tmp_path)
    make_input(case["input"], os.path.join(d, "in.fastq"))
    a = ["0", "0", "1"] if case["mode"] == "il" else ["0", "1", "0"]
    subprocess.run([os.path.join(BIN, "genFastqIdx"), str(case["C"])] + a + ["in.fastq"], cwd=d, check=True,
                   stdout=subprocess.DEVNULL)
    assert masked_sha(os.path.join(d, "in.fastq.idx")) == case["sha256"]["idx"]


@pytest.mark.parametrize("case", [c for c in golden()["cases"]], ids=[c["name"] for c in golden()["cases"]])
def test_view_index_tool_matches_reference(case, tmp_path):
    """viewIdx (viewIdx.c:18-49) over this build's .idx prints what the
    reference's viewIdx prints over the reference's .idx (whose uninitialised
    bytes are zeroed, as this build writes them)."""
    import hashlib
    d = str(tmp_path)
    make_input(case["input"], os.path.join(d, "in.fastq"))
    a = ["0", "0", "1"] if case["mode"] == "il" else ["0", "1", "0"]
    subprocess.run([os.path.join(BIN, "genFastqIdx"), str(case["C"])] + a + ["in.fastq"], cwd=d, check=True,
                   stdout=subprocess.DEVNULL)
    r = subprocess.run([os.path.join(BIN, "viewIdx"), "in.fastq.idx"], cwd=d, check=True, stdout=subprocess.PIPE)
    assert hashlib.sha256(r.stdout).hexdigest() == case["viewidx_sha256"]
    lines = r.stdout.decode().splitlines()
    assert len(lines) == case["C"] + 2 and lines[-1] == "TotalKmers 0"


def test_two_file_paired_end_input_is_refused(tmp_path):
    """Two-file paired-end input is refused with an error by every tool and by
    the C ABI: the reference gives both mates one read id (metapart.c:527-540),
    sizes its buffers from histograms that never read the mate file
    (genHistograms.c reads files[0] only), and its writer drops those reads
    (metapart.c:1037-1047), so there is no defined result to reproduce."""
    d = str(tmp_path)
    buf = mp.synth(0, n_reads=200)
    for f in ("a_1.fastq", "a_2.fastq"):
        open(os.path.join(d, f), "wb").write(buf.tobytes())
    os.makedirs(os.path.join(d, "out", "Rank0"))
    files = ["a_1.fastq", "a_2.fastq"]
    for cmd in ([os.path.join(BIN, "genFastqIdx"), "2", "1", "0", "0"] + files,
                [os.path.join(BIN, "genHistograms_metapart"), "27", "1", "0", "0"] + files,
                [os.path.join(BIN, "genHistograms_metapartmin"), "27", "1", "0", "0"] + files,
                [os.path.join(BIN, "metapart"), "-o", "./out", "2000000000", "27", "1", "1", "1", "0", "0"] + files,
                [os.path.join(BIN, "metapartmin"), "-o", "./out", "2000000000", "27", "1", "1", "1", "0", "0"] + files):
        r = subprocess.run(cmd, cwd=d, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
        assert r.returncode != 0, cmd[0]
        assert "not supported" in r.stderr or "unsupported" in r.stderr, (cmd[0], r.stderr)
    assert not os.path.exists(os.path.join(d, "a_1.fastq.idx"))
    assert os.listdir(os.path.join(d, "out", "Rank0")) == []
    arr = (ct.c_char_p * 2)(*[os.path.join(d, f).encode() for f in files])
    assert mp.lib().mp_index_fastq(2, 1, 0, 0, arr, 2) == -1


def test_index_tool_refuses_empty_input(tmp_path):
    """An empty FASTQ, on which the reference's genFastqIdx stops at an
    assertion (genFastqIdx.c:250-251: opening a chunk's file fails): the tool
    exits with an error and leaves no index behind."""
    d = str(tmp_path)
    open(os.path.join(d, "in.fastq"), "wb").close()
    r = subprocess.run([os.path.join(BIN, "genFastqIdx"), "4", "0", "1", "0", "in.fastq"], cwd=d,
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    assert r.returncode != 0 and "error" in r.stderr
    assert not os.path.exists(os.path.join(d, "in.fastq.idx"))


def test_index_tool_refuses_records_past_the_boundary_window(tmp_path):
    """Reads of 2-4 kb: a chunk boundary must fall in the 3000 bytes the
    reference's walk reads (genFastqIdx.c:254), and with records longer than
    that its genFastqIdx aborts (the assertion at genFastqIdx.c:304).  The tool exits with an error and leaves no
    index behind; the oracle names the chunk."""
    d = str(tmp_path)
    buf = mp.synth(0, n_reads=400, n_genomes=2, genome_len_min=40000, genome_len_max=60000, read_len_min=2000,
                   read_len_max=4000, err_rate=0.01, seed=8)
    open(os.path.join(d, "in.fastq"), "wb").write(buf.tobytes())
    r = subprocess.run([os.path.join(BIN, "genFastqIdx"), "8", "0", "1", "0", "in.fastq"], cwd=d,
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    assert r.returncode != 0 and "error" in r.stderr
    assert not os.path.exists(os.path.join(d, "in.fastq.idx"))
    r = subprocess.run([ORACLE_CLI, "genFastqIdx", "8", "0", "1", "0", "in.fastq"], cwd=d,
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    assert r.returncode != 0 and "no record boundary" in r.stderr


def test_plan_ranges_refuses_an_empty_histogram():
    """No k-mers at all: the reference's planning walk writes past its arrays
    (its metapart crashes on such an input); mp_plan_ranges refuses."""
    with pytest.raises(mp.MPError) as e:
        mp.plan_ranges(np.zeros(2 ** 20 + 1, dtype=np.uint64), 1, 1)
    assert e.value.code == -3


def test_index_buffer_equals_index_file(tmp_path):
    buf = mp.synth(1)
    p = os.path.join(str(tmp_path), "in.fastq")
    open(p, "wb").write(buf.tobytes())
    subprocess.run([os.path.join(BIN, "genFastqIdx"), "8", "0", "1", "0", "in.fastq"], cwd=str(tmp_path), check=True,
                   stdout=subprocess.DEVNULL)
    raw = np.fromfile(p + ".idx", dtype=np.uint8).reshape(-1, mp.FILEINDEX_BYTES)
    offs = raw[:, 0:8].copy().view(np.uint64).ravel()
    assert np.array_equal(mp.index_buffer(buf, 8), offs)


def test_plan_ranges_matches_oracle():
    rng = np.random.default_rng(11)
    L = oracle()
    for trial in range(6):
        counts = rng.poisson(rng.uniform(0.5, 40), size=2 ** 20).astype(np.uint64)
        H = np.zeros(2 ** 20 + 1, dtype=np.uint64)
        H[1:] = np.cumsum(counts)
        for S, T in [(1, 1), (2, 2), (4, 3), (7, 1)]:
            scan = np.zeros(S + 1, dtype=np.uint64)
            thr = np.zeros(S * T + 1, dtype=np.uint64)
            rc = L.orc_plan_ranges(H.ctypes.data, S, T, scan.ctypes.data, thr.ctypes.data)
            assert rc == 0
            s2, t2 = mp.plan_ranges(H, S, T)
            assert np.array_equal(scan, s2) and np.array_equal(thr, t2)


def test_plan_ranges_rejects_degenerate():
    H = np.zeros(2 ** 20 + 1, dtype=np.uint64)
    H[1:] = 5  # every k-mer in bucket 0: no thread split is possible
    with pytest.raises(mp.MPError) as e:
        mp.plan_ranges(H, 1, 4)
    assert e.value.code == -3


def test_single_scan_plan_covers_every_tuple():
    """With one scan and one segment the plan is [0, last non-empty bucket + 1):
    a batch that enumerates every bucket of a sample sees exactly the tuples of
    that sample's plain run."""
    rng = np.random.default_rng(5)
    for trial in range(8):
        counts = (rng.random(2 ** 20) < rng.uniform(0.001, 0.9)).astype(np.uint64) * rng.integers(1, 9, 2 ** 20).astype(np.uint64)
        if trial % 2:
            counts[-rng.integers(1, 5000):] = 0
        if counts.sum() == 0:
            counts[rng.integers(0, 2 ** 20)] = 1
        H = np.zeros(2 ** 20 + 1, dtype=np.uint64)
        H[1:] = np.cumsum(counts)
        scan, thr = mp.plan_ranges(H, 1, 1)
        last = int(np.nonzero(counts)[0][-1])
        assert int(scan[0]) == 0 and int(scan[1]) == last + 1
        assert list(thr) == list(scan)
