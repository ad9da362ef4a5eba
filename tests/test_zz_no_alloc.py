"""The allocation-free histogram entry point (mp_count_chunks_noalloc, C ABI)
gives the per-chunk m-mer counts, reads per chunk and k-mer total of
mp_count_chunks on the same FASTQ bytes, on two calls in a row, the second
with a smaller input (the fixed scratch keeps the first call's contents), and
rejects an input above its size limit."""
import ctypes as ct

import numpy as np
import pytest

import metapart_amd as mp


def _device_fastq(buf):
    import torch
    host = np.frombuffer(buf.tobytes(), dtype=np.uint8)
    d = torch.empty(host.size + 64, dtype=torch.uint8, device="cuda")
    d[: host.size].copy_(torch.from_numpy(host.copy()))
    torch.cuda.synchronize()
    return d


def _noalloc(d, nbytes, cofs, k, min3):
    C = len(cofs) - 1
    counts = np.zeros((C, mp.NBUCKETS + 2), dtype=np.uint32)
    reads = np.zeros(C, dtype=np.uint64)
    tot = ct.c_uint64(0)
    co = np.ascontiguousarray(np.asarray(cofs, dtype=np.uint64))
    rc = mp.lib().mp_count_chunks_noalloc(0, k, 1 if min3 else 0, ct.c_void_p(d.data_ptr()), nbytes,
                                         co.ctypes.data_as(ct.POINTER(ct.c_uint64)), C, counts.ctypes.data,
                                         reads.ctypes.data_as(ct.POINTER(ct.c_uint64)), ct.byref(tot))
    return rc, counts, reads, int(tot.value)


def _reference(buf, cofs, k, min3):
    C = len(cofs) - 1
    counts = np.zeros((C, mp.NBUCKETS + 2), dtype=np.uint32)
    reads = np.zeros(C, dtype=np.uint64)
    tot = ct.c_uint64(0)
    co = np.ascontiguousarray(np.asarray(cofs, dtype=np.uint64))
    with mp.Context(k=k, min3=min3) as ctx:
        ptr, n = mp._buf_ptr(buf)
        ctx._check(mp.lib().mp_count_chunks(ctx._h, ptr, n, co.ctypes.data_as(ct.POINTER(ct.c_uint64)), C,
                                            counts.ctypes.data, reads.ctypes.data_as(ct.POINTER(ct.c_uint64)),
                                            ct.byref(tot)), "mp_count_chunks")
    return counts, reads, int(tot.value)


@pytest.mark.gpu
@pytest.mark.parametrize("min3", [False, True], ids=["plain", "min3"])
def test_noalloc_counts_equal_count_chunks(min3):
    print("gpu no-alloc histogram, %s" % ("3-minimizer" if min3 else "plain"), flush=True)
    big = mp.synth(0, n_reads=60000, read_len=150, n_genomes=20, genome_len_min=20000, genome_len_max=60000,
                   seed=71, n_rate=0.001)
    small = mp.synth(0, n_reads=9000, read_len_min=40, read_len_max=140, seed=72)
    for buf, C, k in ((big, 8, 31), (small, 3, 27)):
        cofs = mp.index_buffer(buf, C)
        d = _device_fastq(buf)
        rc, counts, reads, tot = _noalloc(d, buf.nbytes, cofs, k, min3)
        assert rc == 0
        want = _reference(buf, cofs, k, min3)
        assert np.array_equal(counts, want[0])
        assert np.array_equal(reads, want[1])
        assert tot == want[2] > 0
        del d


@pytest.mark.gpu
def test_noalloc_rejects_inputs_beyond_its_limits():
    print("gpu no-alloc histogram: limits", flush=True)
    buf = mp.synth(0, n_reads=2000)
    d = _device_fastq(buf)
    cofs = mp.index_buffer(buf, 2)
    rc, _, _, _ = _noalloc(d, (32 << 20) + 1, [0, (32 << 20) + 1], 27, False)
    assert rc == -1
    rc, _, _, _ = _noalloc(d, buf.nbytes, np.linspace(0, buf.nbytes, 10).astype(np.uint64), 27, False)
    assert rc == -1  # more than 8 chunks
    rc, counts, _, tot = _noalloc(d, buf.nbytes, cofs, 27, False)
    assert rc == 0 and tot == int(counts.sum()) > 0
